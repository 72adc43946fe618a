"""Benchmark of the spray droplet update on one MI355X.

Workload (config 2): 128^3 periodic box, dx = 1e-4 m, air at 30 atm,
T0 = 900 K with the sinusoidal recipe, ~1.6e7 droplets (Poisson ~7.63 per
cell) of an n-decane/n-dodecane liquid, median diameter 15 um, sigma_ln 0.5,
drag + evaporation + full momentum/mass/energy deposition, dt = 5e-7 s,
dtmod = 1 (one call = kick + drift).  A step is one updateParticles call
over all droplets that leaves the step's source in the FAB: the update runs
in the source overwrite mode, which writes what the reference's zero fill
(transferSource's tmpSource.setVal(0), SprayParticles.H:294) followed by
the accumulating update gives.

Prints one JSON line.  `value` is the FP64 (reference method) throughput in
droplet-steps/s; the FP32 form is reported under "forms".  The roofline entry
is for the dominant kernel (the fused update kernel of the FP64 form):
achieved = algorithmic bytes per droplet-step x droplets per launch / mean
kernel time from HIP events on the launch stream.  cpu_baseline times the
CPU oracle (the FP64 restatement of the reference method) on a bounded
sample of the same workload on this host's cores.

    python bench.py [--gpus 1] [--steps 50] [--warmup 2] [--forms fp64,fp32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "droplet-steps/s at 16M drops, plain FP64 vs MLP-surrogate; % of HBM peak"
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6


def algorithmic_bytes(nf: int, ns: int, ppc: float, w: int) -> float:
    """Compulsory HBM bytes per droplet-step: droplet reads (9+nf) words and
    writes (8+nf) words of width w; per cell, the gas state (5+NS) FP64 words
    are read and the sources (5+nf) FP64 words written, shared by ppc
    droplets."""
    return (17 + 2 * nf) * w + (5 + ns + 5 + nf) * 8.0 / ppc


def run_form(prob, form: int, steps: int, warmup: int, stream, acc_steps: int = 0):
    import ctypes

    from pelemp_amd import _capi as C
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    cap = int(prob.n * 1.02) + 1024
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, cap, form=form)
    net = None
    if form == C.FORM_SURROGATE:
        from pelemp_amd import surrogate as S

        net = S.Surrogate.load(S.default_path(prob.mech))
        net.install(pc)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    src = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
    h = stream.handle

    # the update writes the source's cells (the caller's setVal(0.) of the
    # reference, SprayParticles.H:294, fused into the update)
    pc.set_source_mode(C.SOURCE_OVERWRITE)

    def step():
        pc.updateParticles(0, state, src, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                           True, None, -1.0, stream=h, sync=False)

    for _ in range(warmup):
        step()
    pc.sync(h)
    pc.pass_counts()  # reset the count log
    pc.set_timing(True)
    timer = C.Timer()
    C.device_synchronize()
    t0 = time.perf_counter()
    timer.start(h)
    for _ in range(steps):
        step()
    timer.stop(h)
    stream.synchronize()
    C.device_synchronize()
    wall = time.perf_counter() - t0
    pc.sync(h)
    ms_total = timer.elapsed_ms()
    counts = pc.pass_counts()
    kms, kn = pc.get_timing()
    pc.set_timing(False)
    droplet_steps = int(sum(counts))
    res = dict(droplet_steps=droplet_steps, ms_total=ms_total, wall_s=wall, counts=counts,
               update_ms=kms[0] / max(kn[0], 1), update_launches=kn[0], sort_ms=kms[1] / max(kn[1], 1),
               rate=droplet_steps / (ms_total * 1e-3), n_end=pc.num_particles())
    if net is not None:
        ns, nfb = ctypes.c_int64(), ctypes.c_int64()
        C.check(C.lib().pelemp_spray_surrogate_counts(pc._h, ctypes.byref(ns), ctypes.byref(nfb)), "counts")
        res["surrogate_share"] = ns.value / max(1, ns.value + nfb.value)
    if acc_steps > 0:
        # untimed accuracy pass from the initial droplets: evaporated mass =
        # sum over steps of the mass source (accumulated, += semantics) * V * dt
        pc.set_particles(prob.pos, prob.rdata, prob.ids)
        pc.set_source_mode(C.SOURCE_ACCUMULATE)
        src.setVal(0.0)
        for _ in range(acc_steps):
            pc.updateParticles(0, state, src, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                               True, None, -1.0, stream=h, sync=False)
        pc.sync(h)
        V = float(np.prod(prob.geom["dx"]))
        res["evaporated_mass"] = float(src.numpy()[prob.comps["rhoSrcIndx"]].sum()) * V * prob.flow_dt
        res["acc_steps"] = acc_steps
    pc.close()
    return res


def _liquid_mass(prob, rd) -> float:
    """Liquid mass of droplets in the AoS meaning: N pi/6 rho_L d^3 with
    rho_L the mass-fraction-weighted harmonic mean of the fuels' cubic fits
    at T_d (SprayFuelData.H:146-155)."""
    nf = prob.nf
    c = np.asarray(prob.sdata["rho_coef"], dtype=np.float64)
    T = rd[:, 3]
    rinv = np.zeros(rd.shape[0])
    for k in range(nf):
        a, b, cc, d = c[4 * k:4 * k + 4]
        rinv += rd[:, 5 + k] / (a + T * (b + T * (cc + T * d)))
    return float(np.sum(rd[:, 5 + nf] * np.pi / 6.0 / rinv * rd[:, 4] ** 3))


def lifetime_study(prob, stream, forms=("fp64", "fp32", "surrogate"), max_steps: int = 6000, every: int = 10,
                   field_steps=(100, 500, 1000)):
    """Errors of the FP32 and surrogate forms against the FP64 form (which
    equals the oracle bit for bit) over whole droplet lifetimes, from the same
    droplets in the same frozen gas: each form runs until its liquid is gone
    (or max_steps).  Every `every` steps the droplets are sampled; a droplet
    that vanished since the last sample gets the lifetime its last two
    samples give by the d^2 law.  Reported per form: the times at which 50,
    90 and 99 % of the initial liquid mass has evaporated, the per-droplet
    lifetime error against the FP64 form (median, p99, max of |dt|/t), the
    largest difference of the evaporated-mass curves as a fraction of the
    initial mass, and, at the steps in field_steps, the error of the
    accumulated gas-phase source fields per component: the largest cell
    difference relative to the FP64 field's largest value, and the relative
    L2 difference.  A droplet that crosses a cell face one step earlier or
    later moves that step's deposit to the neighbouring cell, which the
    largest cell difference shows and the evaporated mass does not."""
    from pelemp_amd import _capi as C
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    h = stream.handle
    V = float(np.prod(prob.geom["dx"]))
    dt = prob.flow_dt
    m0 = _liquid_mass(prob, prob.rdata)
    nid = int(prob.ids.max()) + 1
    code = {"fp64": C.FORM_FP64, "fp32": C.FORM_FP32, "surrogate": C.FORM_SURROGATE}
    runs = {}
    for name in forms:
        pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, int(prob.n * 1.02) + 1024,
                                    form=code[name])
        if name == "surrogate":
            from pelemp_amd import surrogate as S

            S.Surrogate.load(S.default_path(prob.mech)).install(pc)
        pc.set_particles(prob.pos, prob.rdata, prob.ids)
        state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
        src = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
        pc.set_source_mode(C.SOURCE_ACCUMULATE)
        life = np.full(nid, np.nan)
        d2a = np.full(nid, np.nan)  # d^2 at the last sample
        d2b = np.full(nid, np.nan)  # d^2 at the sample before
        d2a[prob.ids] = prob.rdata[:, 4] ** 2
        ta, tb = 0.0, np.nan
        alive = np.zeros(nid, bool)
        alive[prob.ids] = True
        curve_t, curve_m = [0.0], [0.0]
        fields = {}
        step = 0
        while step < max_steps and alive.any():
            for _ in range(every):
                pc.updateParticles(0, state, src, dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False, True,
                                   None, -1.0, stream=h, sync=False)
            step += every
            pc.sync(h)
            t = step * dt
            s = src.numpy()
            curve_t.append(t)
            curve_m.append(float(s[prob.comps["rhoSrcIndx"]].sum()) * V * dt)
            if step in field_steps:
                fields[step] = s.copy()
            _, rd, ids = pc.get_particles()
            now = np.zeros(nid, bool)
            now[ids] = True
            gone = np.flatnonzero(alive & ~now)
            if gone.size:
                slope = (d2a[gone] - d2b[gone]) / (ta - tb) if np.isfinite(tb) else np.full(gone.size, np.nan)
                est = ta - d2a[gone] / np.where(slope < 0.0, slope, np.nan)
                life[gone] = np.clip(np.where(np.isfinite(est), est, t), ta, t)
            alive = now
            d2b[ids] = d2a[ids]
            d2a[ids] = rd[:, 4] ** 2
            tb, ta = ta, t
        if name == "surrogate":
            import ctypes

            ns, nfb = ctypes.c_int64(), ctypes.c_int64()
            C.check(C.lib().pelemp_spray_surrogate_counts(pc._h, ctypes.byref(ns), ctypes.byref(nfb)), "counts")
            share = ns.value / max(1, ns.value + nfb.value)
        else:
            share = None
        pc.close()
        runs[name] = dict(life=life, t=np.array(curve_t), m=np.array(curve_m) / m0, fields=fields, steps=step,
                          left=int(alive.sum()), share=share)

    def t_at(r, frac):
        k = np.flatnonzero(r["m"] >= frac)
        return float(r["t"][k[0]]) * 1e3 if k.size else None

    ref = runs.get("fp64")
    out = dict(droplets=int(prob.n), dt=dt, every=every, initial_liquid_mass=m0)
    for name, r in runs.items():
        o = dict(steps=r["steps"], droplets_left=r["left"], evaporated_frac_end=float(r["m"][-1]),
                 t50_ms=t_at(r, 0.5), t90_ms=t_at(r, 0.9), t99_ms=t_at(r, 0.99))
        if r["share"] is not None:
            o["surrogate_share"] = r["share"]
        if ref is not None and name != "fp64":
            both = np.isfinite(r["life"]) & np.isfinite(ref["life"])
            rel = np.abs(r["life"][both] - ref["life"][both]) / ref["life"][both]
            o["lifetime_droplets_compared"] = int(both.sum())
            if rel.size:
                o["lifetime_rel_err"] = dict(median=float(np.median(rel)), p99=float(np.quantile(rel, 0.99)),
                                             max=float(rel.max()))
            k = min(len(r["m"]), len(ref["m"]))
            o["evaporated_frac_max_abs_err"] = float(np.max(np.abs(r["m"][:k] - ref["m"][:k])))
            ferr = {}
            for st_, f in r["fields"].items():
                g = ref["fields"].get(st_)
                if g is None:
                    continue
                scale = np.max(np.abs(g), axis=(1, 2, 3))
                err = np.max(np.abs(f - g), axis=(1, 2, 3)) / np.where(scale > 0, scale, 1.0)
                l2g = np.sqrt(np.sum(g * g, axis=(1, 2, 3)))
                l2 = np.sqrt(np.sum((f - g) ** 2, axis=(1, 2, 3))) / np.where(l2g > 0, l2g, 1.0)
                nfp = prob.nf
                ferr[str(st_)] = dict(max=dict(mom=float(err[0:3].max()), mass=float(err[3]),
                                               species=float(err[4:4 + nfp].max()), energy=float(err[4 + nfp])),
                                      l2=dict(mom=float(l2[0:3].max()), mass=float(l2[3]),
                                              species=float(l2[4:4 + nfp].max()), energy=float(l2[4 + nfp])))
            o["source_field_err"] = ferr
        out[name] = o
    return out


def run_lifetimes(stream, ncell: int = 16):
    """The lifetime study on the config-4 physics (1 atm, 1000 K, n-dodecane,
    dt 1e-6 s) and the config-2 physics (30 atm, 900 K, n-decane/n-dodecane,
    dt 5e-7 s), each on a small box with the workload's droplet density."""
    from pelemp_amd import problem as P

    out = {}
    p4 = P.config4(ncell=ncell)
    out["config4"] = dict(workload=f"config4 recipe on {ncell}^3 cells",
                          **lifetime_study(p4, stream, max_steps=16000, every=10, field_steps=(100, 1000, 4000)))
    p2 = P.config2(ncell=ncell)
    out["config2"] = dict(workload=f"config2 recipe on {ncell}^3 cells",
                          **lifetime_study(p2, stream, max_steps=32000, every=20, field_steps=(200, 2000, 8000)))
    return out


def run_config4(steps: int, warmup: int, stream, acc_steps: int = 10):
    """Config 4: the learned surrogate on the config-1 physics scaled to
    1.6e7 droplets (128^3 cells, 1 atm, 1000 K, n-dodecane), beside the FP64
    and FP32 forms on the same droplets: droplet-steps/s, the share of
    droplet-substeps the network served, and the evaporated mass of
    acc_steps steps against the FP64 form (errors over whole lifetimes are
    in the lifetimes entry)."""
    from pelemp_amd import _capi as C
    from pelemp_amd import problem as P

    t0 = time.perf_counter()
    prob = P.config4()
    print(f"# config4 built: {prob.n} droplets in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    out = dict(workload=f"config4: 128^3 periodic box dx=1e-4 m, 1 atm, T0=1000 K, {prob.n} n-dodecane droplets "
                        f"(Poisson ~7.63/cell, median 20 um, sigma_ln 0.4, T_d 300 K), dt=1e-6 s",
               droplets=int(prob.n), forms={})
    code = {"fp64": C.FORM_FP64, "fp32": C.FORM_FP32, "surrogate": C.FORM_SURROGATE}
    for name in ("fp64", "fp32", "surrogate"):
        r = run_form(prob, code[name], steps, warmup, stream, acc_steps)
        f = dict(value=r["rate"], ms_per_step=r["ms_total"] / steps, update_kernel_ms=r["update_ms"],
                 sort_ms=r["sort_ms"], evaporated_mass=r.get("evaporated_mass"))
        if "surrogate_share" in r:
            f["surrogate_share"] = r["surrogate_share"]
        out["forms"][name] = f
        print(f"# config4 {name}: {r['rate']:.4g} droplet-steps/s, {f['ms_per_step']:.3f} ms/step, "
              f"update kernel {r['update_ms']:.3f} ms", file=sys.stderr, flush=True)
    ref = out["forms"]["fp64"]
    for name, f in out["forms"].items():
        f["speedup_vs_fp64"] = f["value"] / ref["value"]
        if ref.get("evaporated_mass"):
            f["evaporated_mass_rel_err_vs_fp64"] = abs(f["evaporated_mass"] - ref["evaporated_mass"]) / abs(
                ref["evaporated_mass"])
    return out


def run_branches(steps: int, warmup: int, stream, cap_factor: float = 5.0):
    """Config 3 (FP64 with walls, TAB breakup, splash and wall films): the
    update kernel with its branches inline, plus the passes around it -- the
    film-thickness pass before the update and the secondary-droplet pass
    (child counts, scan, creation) after it.  Their time is the
    branch-compaction cost.  The parcel count grows with the breakup cascade
    (about 3.6x over the first 6 steps), so the leg runs 1 + 5 steps with
    room for 5x the initial droplets."""
    from pelemp_amd import _capi as C
    from pelemp_amd import problem as P
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    t0 = time.perf_counter()
    prob = P.config3()
    print(f"# config3 built: {prob.n} droplets in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    cap = int(prob.n * cap_factor)
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, cap, form=C.FORM_FP64)
    pc.set_params(100.0, 0.5, 1)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    src = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
    h = stream.handle
    n0 = prob.n

    # the update writes the source's cells (the caller's setVal(0.) of the
    # reference, SprayParticles.H:294, fused into the update)
    pc.set_source_mode(C.SOURCE_OVERWRITE)

    def step():
        pc.updateParticles(0, state, src, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                           True, None, -1.0, stream=h, sync=False)

    for _ in range(warmup):
        step()
    pc.sync(h)
    n_start = pc.num_particles()
    pc.pass_counts()
    pc.set_timing(True)
    timer = C.Timer()
    C.device_synchronize()
    timer.start(h)
    for _ in range(steps):
        step()
    timer.stop(h)
    stream.synchronize()
    pc.sync(h)
    ms_total = timer.elapsed_ms()
    counts = pc.pass_counts()
    kms, kn = pc.get_timing()
    n_end = pc.num_particles()
    pc.close()
    ds = int(sum(counts))
    per = [kms[i] / max(kn[i], 1) for i in range(4)]
    branch_ms = per[2] + per[3]
    return dict(workload=f"config3: 128x128x256 slab dx=1e-4 m, wall at z=0 (reflective, T_wall=450 K), "
                         f"outflow at z_hi, gas 600 K / 5 atm moving at -20 m/s, {n0} n-dodecane droplets "
                         "(median 40 um, sigma_ln 0.6, We 1-300), TAB + splash + wall films, FP64, dt=5e-7 s",
                droplets_start=n0, droplets_timed_start=n_start, droplets_end=n_end, droplet_steps=ds,
                value=ds / (ms_total * 1e-3), ms_per_step=ms_total / steps, update_kernel_ms=per[0],
                sort_ms=per[1], film_pass_ms=per[2], children_pass_ms=per[3], branch_compaction_ms=branch_ms,
                branch_compaction_frac=branch_ms * steps / ms_total)


def run_jet(steps: int, warmup: int, stream, n_drops: int, ncell: int = 256, forms=("fp32", "fp64")):
    """Config 5 (dense uneven jet): ncell^3 box at 60 atm / 800 K, n_drops
    droplets of a hollow-cone jet with a six-component alkane liquid, the
    gas state in FP32, 20-step runs at dt = 2e-7 s.  Reports per form the
    droplet-steps/s and, from the droplets' cells, the load of the update
    workgroups: droplets per tile (one workgroup per tile) and per work item
    once tiles above PELEMP_SEGMENT_DROPLETS are split into segments."""
    from pelemp_amd import _capi as C
    from pelemp_amd import problem as P
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    t0 = time.perf_counter()
    prob = P.config5(ncell=ncell, n_drops=n_drops, droplets=False)
    pos, rd, ids = P.jet_droplets(prob, n_drops)
    gen_s = time.perf_counter() - t0
    print(f"# config5 built: {n_drops} droplets on {ncell}^3 cells in {gen_s:.1f} s", file=sys.stderr, flush=True)
    state32 = prob.state.astype(np.float32)
    # workgroup load from the droplets' tiles
    dx = prob.geom["dx"][0]
    nt = ncell // 4
    c = np.floor(pos * (1.0 / dx)).astype(np.int64)
    np.clip(c, 0, ncell - 1, out=c)
    tiles = np.bincount(((c[:, 2] // 4) * nt + c[:, 1] // 4) * nt + c[:, 0] // 4, minlength=nt ** 3)
    cells = np.bincount((c[:, 2] * ncell + c[:, 1]) * ncell + c[:, 0], minlength=ncell ** 3)
    del c
    seg = C.SEGMENT_DROPLETS
    occ = tiles[tiles > 0]
    nseg = (occ + seg - 1) // seg
    items = np.concatenate([np.minimum(seg, occ - k * seg) for k in range(int(nseg.max()))])
    items = items[items > 0]
    out = dict(workload=f"config5: {ncell}^3 periodic box dx=1e-4 m, 60 atm / 800 K gas in FP32, {n_drops} droplets "
                        "of a hollow-cone jet (10 deg +- 3 deg, 200 m/s +- 10 %), six C8-C16 n-alkanes "
                        "(Dirichlet(2)), median 10 um, dt=2e-7 s",
               droplets=n_drops, generate_s=gen_s, cells_occupied=int((cells > 0).sum()),
               max_per_cell=int(cells.max()), cells_over_500=int((cells > 500).sum()),
               tiles_occupied=int(occ.size), max_per_tile=int(occ.max()), mean_per_tile=float(occ.mean()),
               tile_imbalance=float(occ.max() / occ.mean()), work_items=int(items.size),
               max_per_item=int(items.max()), mean_per_item=float(items.mean()),
               item_imbalance=float(items.max() / items.mean()), forms={})
    del cells
    for name in forms:
        form = {"fp64": C.FORM_FP64, "fp32": C.FORM_FP32}[name]
        pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, int(n_drops * 1.01) + 1024,
                                    form=form)
        t1 = time.perf_counter()
        pc.set_particles(pos, rd, ids)
        set_s = time.perf_counter() - t1
        state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], state32, dtype=np.float32)
        src = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
        h = stream.handle

        pc.set_source_mode(C.SOURCE_OVERWRITE)

        def step():
            pc.updateParticles(0, state, src, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False,
                               False, True, None, -1.0, stream=h, sync=False)

        for _ in range(warmup):
            step()
        pc.sync(h)
        pc.pass_counts()
        pc.set_timing(True)
        timer = C.Timer()
        C.device_synchronize()
        timer.start(h)
        for _ in range(steps):
            step()
        timer.stop(h)
        stream.synchronize()
        pc.sync(h)
        ms = timer.elapsed_ms()
        counts = pc.pass_counts()
        kms, kn = pc.get_timing()
        pc.set_timing(False)
        ds = int(sum(counts))
        out["forms"][name] = dict(value=ds / (ms * 1e-3), ms_per_step=ms / steps, update_kernel_ms=kms[0] / max(kn[0], 1),
                                  sort_ms=kms[1] / max(kn[1], 1), droplet_steps=ds, droplets_end=pc.num_particles(),
                                  set_particles_s=set_s)
        print(f"# config5 {name}: {ds / (ms * 1e-3):.4g} droplet-steps/s, {ms / steps:.3f} ms/step, update kernel "
              f"{kms[0] / max(kn[0], 1):.3f} ms", file=sys.stderr, flush=True)
        pc.close()
        del state, src
    if "fp64" in out["forms"] and "fp32" in out["forms"]:
        out["fp32_speedup_vs_fp64"] = out["forms"]["fp32"]["value"] / out["forms"]["fp64"]["value"]
    return out


def host_cpu() -> dict:
    """The host the CPU baseline runs on: logical CPUs, the CPUs this process
    may run on, the OpenMP thread setting and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    lscpu = {}
    try:
        import subprocess

        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20)
        keep = ("Architecture", "CPU(s)", "On-line CPU(s) list", "Model name", "Thread(s) per core",
                "Core(s) per socket", "Socket(s)", "NUMA node(s)", "CPU max MHz", "L3 cache")
        for line in r.stdout.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keep:
                lscpu[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    compiler = None
    try:
        from oracle import oracle_py as O

        compiler = O.compiler() + " -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (oracle/Makefile)"
    except (OSError, AttributeError):
        pass
    return dict(cpu_count=os.cpu_count(), affinity=aff, omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                model=model, nproc=aff, lscpu=lscpu, compiler=compiler)


def cpu_baseline(prob, target_s: float = 12.0, parity_stream=None):
    """Time the CPU oracle on a bounded sample of the workload, on every CPU
    the process may use (the OpenMP thread count, when set, is the CPU share
    the job is given).  When the sample is the whole droplet set, the GPU
    FP64 form runs the same steps from the same droplets and is compared with
    the oracle bit for bit (full-size parity)."""
    from oracle import oracle_py as O

    host = host_cpu()
    nthreads = host["affinity"]
    if host["omp_num_threads"]:
        nthreads = min(nthreads, int(host["omp_num_threads"]))
    sd = O.latent_shifted(prob.sdata, prob.mech, len(prob.species))

    def timed(stride: int, nsteps: int, keep=False):
        pos, rd, ids = prob.pos[::stride], prob.rdata[::stride], prob.ids[::stride]
        t0 = time.perf_counter()
        nsum = 0
        srcs = []
        for _ in range(nsteps):
            src = prob.new_source()
            nsum += ids.shape[0]
            st, pos, rd, ids = O.update(prob.mech, sd, prob.comps, prob.geom, pos, rd, ids, prob.state,
                                        prob.state_lo, prob.state_hi, src, prob.src_lo, prob.src_hi, prob.flow_dt,
                                        nthreads=nthreads, state_ghosts=prob.state_ghosts,
                                        source_ghosts=prob.source_ghosts)
            assert st == 0
            if keep:
                srcs.append(src)
        return nsum, time.perf_counter() - t0, (pos, rd, ids, srcs)

    # calibrate on ~1/64 of the droplets, then size the sample for ~target_s
    n_cal, t_cal, _ = timed(64, 1)
    per = t_cal / max(n_cal, 1)
    want = max(1, int(target_s / per))
    stride = max(1, int(prob.n // want))
    nsteps = 1
    if stride == 1:
        nsteps = max(1, min(4, int(want // prob.n)))
    n_s, t_s, res = timed(stride, nsteps, keep=(stride == 1 and parity_stream is not None))
    out = dict(value=n_s / t_s, unit="droplet-steps/s", cores=nthreads, kind="port",
               sample=f"{nsteps} step(s) of every {stride}th droplet of the workload "
                      f"({n_s // nsteps} of {prob.n} droplets, full 128^3 gas field), "
                      f"FP64 oracle built with g++ -O3 -march=x86-64-v3 -fopenmp, {t_s:.1f} s",
               seconds=t_s, host=host)
    if stride == 1 and parity_stream is not None:
        out["full_size_parity"] = full_size_parity(prob, nsteps, res)
    return out


def full_size_parity(prob, nsteps: int, oracle_result) -> dict:
    """The GPU FP64 form over the whole workload for nsteps steps, compared
    with the oracle's run of the same steps: droplets by id and every source
    component of every step."""
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    opos, ord_, oids, osrcs = oracle_result
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 1024, form=0)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    src_equal = True
    src_err = 0.0
    for k in range(nsteps):
        src = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
        pc.updateParticles(0, state, src, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                           True, None, -1.0)
        g = src.numpy()
        o = osrcs[k]
        src_equal = src_equal and bool(np.array_equal(g, o))
        scale = np.maximum(np.max(np.abs(o), axis=(1, 2, 3), keepdims=True), 1e-300)
        src_err = max(src_err, float(np.max(np.abs(g - o) / scale)))
        del src
    gpos, grd, gids = pc.get_particles()
    pc.close()
    go, oo = np.argsort(gids), np.argsort(oids)
    same_ids = bool(np.array_equal(gids[go], oids[oo]))
    res = dict(steps=nsteps, droplets=int(prob.n), survivors=int(gids.size), same_droplets=same_ids,
               sources_bitwise=src_equal, max_rel_err_sources=src_err)
    if same_ids:
        res["droplets_bitwise"] = bool(np.array_equal(grd[go], ord_[oo]) and np.array_equal(gpos[go], opos[oo]))
        den = np.maximum(np.abs(ord_[oo]), 1e-3)
        res["max_rel_err_droplets"] = float(np.max(np.abs(grd[go] - ord_[oo]) / den))
    print(f"# full-size parity: {json.dumps(res)}", file=sys.stderr, flush=True)
    return res


def calibrate_copy() -> int:
    """Three device-to-device copies of 1 GiB: the reference kernel for
    scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    from pelemp_amd import _capi as C

    n = 1 << 30
    a, b = C.DeviceBuffer(n), C.DeviceBuffer(n)
    a.zero()
    b.zero()
    for _ in range(3):
        C.check(C.lib().pelemp_memcpy_dtod(b.ptr, a.ptr, n), "copy")
    C.device_synchronize()
    print("copied 3 x 1 GiB", file=sys.stderr)
    return 0


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def measured_traffic(form: str):
    """HBM bytes per launch of the update kernel from the committed rocprofv3
    PMC summary (profiles/spray_update_config2_hbm_traffic.md), or None."""
    path = os.path.join(PROFILES, "spray_update_config2_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch[{form}]"):
            return float(line.split("=")[1].split()[0])
    return None


def valu_profile(key: str, form: str = "fp64"):
    """A per-droplet figure of the update kernel from the committed
    rocprofv3 counter summary (profiles/spray_update_config2_valu_counters.md):
    fp64_flop_per_droplet, valu_instr_per_droplet or valu_busy; None if absent."""
    path = os.path.join(PROFILES, "spray_update_config2_valu_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"{key}[{form}]"):
            return float(line.split("=")[1].split()[0])
    return None


def _timed(fn, reps: int, stream):
    """Mean milliseconds of fn() over reps calls, HIP events on `stream`."""
    from pelemp_amd import _capi as C

    fn()
    C.device_synchronize()
    timer = C.Timer()
    timer.start(stream.handle)
    for _ in range(reps):
        fn()
    timer.stop(stream.handle)
    stream.synchronize()
    return timer.elapsed_ms() / reps


def run_rows(prob, stream, reps: int = 5):
    """The steps around the update on the config-2 problem, FP64: transferSource
    (tmp FAB with 2 ghosts into a state-layout FAB), estTimestep,
    computeDerivedVars (17 variables), the PeleC flavour's update step (same
    droplets and gas in CGS), the update step with a planar EB cutting the
    box, device injection (one call of a jet injecting ~2e5 parcels, and the
    oracle's sequential CPU loop for the same call) and uniformSprayInit of
    the same number of droplets."""
    from pelemp_amd import _capi as C
    from pelemp_amd import problem as P
    from pelemp_amd.spray import DeviceFab, SprayJet, SprayParticleContainer

    h = stream.handle
    out = {}
    cap = int(prob.n * 1.05) + 1024
    g = prob.geom
    n = [g["domain_hi"][d] - g["domain_lo"][d] + 1 for d in range(3)]
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, g, cap)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    src = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
    nsrc = prob.src_shape[0]
    tmp = DeviceFab((-2, -2, -2), (n[0] + 1, n[1] + 1, n[2] + 1), nsrc)
    act = DeviceFab((-1, -1, -1), (n[0], n[1], n[2]), 6 + len(prob.species))
    out["transfer_source_ms"] = _timed(lambda: pc.transferSource(2, 0, tmp, act, stream=h, sync=False), reps, stream)
    t0 = time.perf_counter()
    for _ in range(reps):
        pc.estTimestep(0, 0.5, 0.0, stream=h)
    out["est_timestep_ms"] = (time.perf_counter() - t0) * 1e3 / reps
    nvar = pc.NumDeriveVars(1)
    mf = DeviceFab((0, 0, 0), (n[0] - 1, n[1] - 1, n[2] - 1), nvar)
    out["compute_derived_ms"] = _timed(lambda: pc.computeDerivedVars(mf, 0, 0, 1, stream=h, sync=False), reps, stream)
    out["compute_derived_vars"] = nvar

    def upd(c, st):
        c.updateParticles(0, st, src, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False, True,
                          None, -1.0, stream=h, sync=False)

    out["lmex_step_ms"] = _timed(lambda: upd(pc, state), reps, stream)
    pc.close()
    # PeleC flavour: the same problem in CGS (state rho, rho u, rho E, rho Y, T)
    pcg = P.to_pelec(prob, T_guess_factor=1.02)
    pce = SprayParticleContainer(pcg.mech, pcg.sdata, pcg.comps, pcg.geom, cap, flavour="pelec")
    pce.set_particles(pcg.pos, pcg.rdata, pcg.ids)
    state_c = DeviceFab(pcg.state_lo, pcg.state_hi, pcg.state.shape[0], pcg.state)
    out["pelec_step_ms"] = _timed(lambda: upd(pce, state_c), reps, stream)
    pce.close()
    del pcg, state_c
    # a plane EB (solid below 0.3 of the box height, cutting a layer of cells
    # at 0.37 of a cell); the droplets of the fluid side, short of the top
    # face (whose periodic image is the solid), and of a thin layer behind
    # the EB
    L, dz = g["phi"][2], g["dx"][2]
    eb = P.eb_plane(prob, [0.0, 0.0, 0.3 * L + 0.37 * dz], [0.0, 0.0, -1.0], nsub=4)
    d = (prob.pos - eb["point"]) @ eb["normal"]
    keep = (d < 0.3 * dz) & (prob.pos[:, 2] < L - 3.0 * dz)
    pce = SprayParticleContainer(prob.mech, dict(prob.sdata, min_eb_vfrac=0.05), prob.comps, g, cap)
    pce.set_eb(eb)
    pce.set_particles(prob.pos[keep], prob.rdata[keep], prob.ids[keep])
    out["eb_step_ms"] = _timed(lambda: upd(pce, state), reps, stream)
    out["eb_droplets"] = int(np.count_nonzero(keep))
    out["eb_cut_cells"] = int(np.count_nonzero((eb["flags"] & 3) == 1))
    pce.close()
    del eb
    # injection: a jet in an empty container, ~2e5 parcels per call
    pci = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, g, 4 * 200000 + 1024)
    cent = [0.5 * (a + b) for a, b in zip(g["plo"], g["phi"])]
    mk = dict(jet_cent=cent, jet_norm=[0.0, 0.0, 1.0], spread_angle=20.0, jet_dia=2e-4, jet_vel=50.0, mass_flow=2.0,
              jet_temp=300.0, jet_Y=[1.0 / prob.nf] * prob.nf if prob.nf > 1 else None, dist_type="Normal",
              dist_params=(20e-6, 5e-6), phi_swirl=10.0, inject_ppp=3.0, seed=5)
    jet = SprayJet("jet1", pci, **mk)
    oj = C.SprayJetC.from_buffer_copy(jet._j)  # the checked jet before its first injection
    t0 = time.perf_counter()
    n_inj = pci.sprayInjection(0.0, jet, 1e-6)
    out["inject_ms"] = (time.perf_counter() - t0) * 1e3
    out["inject_droplets"] = n_inj
    try:
        from oracle import oracle_py as O

        t0 = time.perf_counter()
        O.inject(prob.sdata, g, oj, 0.0, 1e-6, 1)
        out["inject_cpu_oracle_ms"] = (time.perf_counter() - t0) * 1e3
    except Exception as e:  # the oracle is optional on the GPU box
        out["inject_cpu_oracle_ms"] = None
        print(f"# rows: oracle injection skipped: {e}", file=sys.stderr, flush=True)
    pci.close()
    # uniformSprayInit of ~1.6e7 droplets
    npart = (256, 256, 244)
    pcu = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, g, int(np.prod(npart)) + 1024)
    t0 = time.perf_counter()
    pcu.uniformSprayInit(npart, (1.0, 0.0, 0.0), 20e-6, 300.0, [1.0 / prob.nf] * prob.nf, num_ppp=2.0)
    C.device_synchronize()
    out["uniform_init_ms"] = (time.perf_counter() - t0) * 1e3
    out["uniform_init_droplets"] = int(np.prod(npart))
    pcu.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # default: config 2's own run length (50 steps of 5e-7 s); the timed
    # window then covers the droplets' heat-up and early evaporation, over
    # which the per-step cost of the forms changes
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--forms", default="fp64,fp32,surrogate")
    ap.add_argument("--acc-steps", type=int, default=10,
                    help="untimed steps per form for the evaporated-mass comparison (0: skip)")
    ap.add_argument("--ncell", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-branches", action="store_true", help="skip the config-3 (walls/breakup/splash) leg")
    ap.add_argument("--no-config4", action="store_true", help="skip the config-4 (surrogate, 1.6e7 droplets) leg")
    ap.add_argument("--no-lifetimes", action="store_true",
                    help="skip the FP32/surrogate error study over whole droplet lifetimes")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the leg timing transferSource, estTimestep, derived variables, PeleC, EB, injection")
    ap.add_argument("--no-jet", action="store_true", help="skip the config-5 (dense hollow-cone jet) leg")
    ap.add_argument("--jet-drops", type=float, default=2.0e8, help="droplets of the config-5 leg")
    ap.add_argument("--jet-only", action="store_true", help="run only the config-5 leg (its own JSON line)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the 1 GiB copy kernels used to calibrate PMC byte counters")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by spatial domain decomposition, see DESIGN.md)", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy()

    from pelemp_amd import _capi as C
    from pelemp_amd import problem as P

    if args.jet_only:
        out = run_jet(args.steps, args.warmup, C.Stream(), int(args.jet_drops))
        print(json.dumps(out))
        return 0
    t0 = time.perf_counter()
    prob = P.config2(ncell=args.ncell)
    print(f"# problem built: {prob.n} droplets on {args.ncell}^3 cells in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    ncells = args.ncell ** 3
    ppc = prob.n / ncells
    ns = len(prob.species)
    stream = C.Stream()
    forms = {}
    for name in args.forms.split(","):
        form = {"fp64": C.FORM_FP64, "fp32": C.FORM_FP32, "surrogate": C.FORM_SURROGATE}[name]
        r = run_form(prob, form, args.steps, args.warmup, stream, args.acc_steps)
        w = 8 if form == C.FORM_FP64 else 4
        b = algorithmic_bytes(prob.nf, ns, ppc, w)
        droplets_per_launch = r["droplet_steps"] / max(r["update_launches"], 1)
        achieved = b * droplets_per_launch / (r["update_ms"] * 1e-3) / 1e9
        forms[name] = dict(value=r["rate"], ms_per_step=r["ms_total"] / args.steps, update_kernel_ms=r["update_ms"],
                           sort_ms=r["sort_ms"], bytes_per_droplet_step=b, achieved_GBs=achieved,
                           hbm_frac=achieved / HBM_PEAK_GBS, droplets_end=r["n_end"],
                           droplet_steps=r["droplet_steps"], traffic=measured_traffic(name))
        for k in ("surrogate_share", "evaporated_mass", "acc_steps"):
            if k in r:
                forms[name][k] = r[k]
        print(f"# {name}: {r['rate']:.4g} droplet-steps/s, {r['ms_total'] / args.steps:.3f} ms/step, "
              f"update kernel {r['update_ms']:.3f} ms, sort {r['sort_ms']:.3f} ms, {achieved:.0f} GB/s",
              file=sys.stderr, flush=True)
    if "fp64" in forms and "evaporated_mass" in forms["fp64"]:
        ref = forms["fp64"]["evaporated_mass"]
        for name, f in forms.items():
            if "evaporated_mass" in f:
                f["evaporated_mass_rel_err_vs_fp64"] = abs(f["evaporated_mass"] - ref) / abs(ref)
        for name, f in forms.items():
            f["speedup_vs_fp64"] = f["value"] / forms["fp64"]["value"]
    main_form = "fp64" if "fp64" in forms else next(iter(forms))
    mf = forms[main_form]
    out = {
        "metric": METRIC,
        "value": mf["value"],
        "unit": "droplet-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mf["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_form == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": f"config2: {args.ncell}^3 periodic box dx=1e-4 m, 30 atm, T0=900 K, "
                               f"{prob.n} n-decane/n-dodecane droplets (Poisson {ppc:.2f}/cell), "
                               "drag+evaporation+deposition, dt=5e-7 s, dtmod=1",
                   "droplets": prob.n, "cells": ncells, "form": main_form},
        "roofline": {"bound": "hbm", "achieved": mf["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": mf["hbm_frac"], "traffic": mf["traffic"],
                     "kernel": "spray_update_kernel", "bytes_per_droplet_step": mf["bytes_per_droplet_step"]},
        "forms": forms,
    }
    flop = valu_profile("fp64_flop_per_droplet")
    if "fp64" in forms and flop is not None and args.ncell == 128:
        # FP64 arithmetic of the FP64 update kernel against the FP64 vector
        # peak: the profiled flop count per droplet-step x droplet-steps per
        # launch / the live mean kernel time
        f64 = forms["fp64"]
        tflops = flop * prob.n / (f64["update_kernel_ms"] * 1e-3) / 1e12
        out["valu"] = {"achieved": tflops, "peak": FP64_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": tflops / FP64_VALU_PEAK_TFLOPS, "fp64_flop_per_droplet_step": flop,
                       "valu_instr_per_droplet_step": valu_profile("valu_instr_per_droplet"),
                       "valu_busy": valu_profile("valu_busy"),
                       "source": "profiles/spray_update_config2_valu_counters.md"}
    if not args.no_rows and "fp64" in forms:
        try:
            out["rows"] = run_rows(prob, stream)
            print(f"# rows: {json.dumps(out['rows'])}", file=sys.stderr, flush=True)
        except C.PelempError as e:
            out["rows"] = dict(error=str(e))
            print(f"# rows failed: {e}", file=sys.stderr, flush=True)
    if not args.no_branches:
        try:
            br = run_branches(min(args.steps, 5), min(args.warmup, 1), stream)
        except C.PelempError as e:
            br = dict(error=str(e))
            print(f"# branches (config3) failed: {e}", file=sys.stderr, flush=True)
        out["branches"] = br
    if "error" not in out.get("branches", {"error": 1}):
        print(f"# branches (config3): {br['value']:.4g} droplet-steps/s, {br['ms_per_step']:.3f} ms/step, "
              f"film pass {br['film_pass_ms']:.3f} ms, children pass {br['children_pass_ms']:.3f} ms, "
              f"droplets {br['droplets_start']} -> {br['droplets_end']}", file=sys.stderr, flush=True)
    if not args.no_config4:
        try:
            out["config4"] = run_config4(min(args.steps, 10), min(args.warmup, 2), stream)
        except C.PelempError as e:
            out["config4"] = dict(error=str(e))
            print(f"# config4 failed: {e}", file=sys.stderr, flush=True)
    if not args.no_lifetimes:
        try:
            out["lifetimes"] = lt = run_lifetimes(stream)
            for k, v in lt.items():
                msg = ", ".join(f"{f}: t50 {v[f]['t50_ms']} ms, lifetime err p99 "
                                f"{v[f].get('lifetime_rel_err', {}).get('p99')}, evap err "
                                f"{v[f].get('evaporated_frac_max_abs_err')}" for f in ("fp64", "fp32", "surrogate")
                                if f in v)
                print(f"# lifetimes {k}: {msg}", file=sys.stderr, flush=True)
        except C.PelempError as e:
            out["lifetimes"] = dict(error=str(e))
            print(f"# lifetimes failed: {e}", file=sys.stderr, flush=True)
    if not args.no_jet:
        try:
            out["jet"] = run_jet(min(args.steps, 20), min(args.warmup, 2), stream, int(args.jet_drops))
        except C.PelempError as e:
            out["jet"] = dict(error=str(e))
            print(f"# jet (config5) failed: {e}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        cb = cpu_baseline(prob, args.cpu_seconds, parity_stream=stream if "fp64" in forms else None)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "host")}
        if "full_size_parity" in cb:
            out["full_size_parity"] = cb["full_size_parity"]
        print(f"# cpu baseline: {cb['value']:.4g} droplet-steps/s on {cb['cores']} threads ({cb['seconds']:.1f} s)",
              file=sys.stderr, flush=True)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
