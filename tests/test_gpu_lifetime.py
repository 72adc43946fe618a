"""Errors of the FP32 and surrogate forms against the FP64 form over whole
droplet lifetimes (bench.lifetime_study): the same droplets in the same
frozen gas until (almost) all liquid has evaporated, on the config-4
recipe (1 atm, 1000 K, n-dodecane, the dodecane network) and the config-2
recipe (30 atm, 900 K, n-decane/n-dodecane, the two-fuel network) on 12^3
cells.  The FP64 form equals the CPU oracle bit for bit (test_gpu_parity).

Bars (the north star's: evaporated mass within 0.5 % of the FP64 oracle):
 * FP32: per-droplet lifetimes within 1e-5 (median) and 1e-4 (p99)
   relative, evaporated-mass curve within 1e-5 of the initial liquid mass,
   accumulated mass-source field within 1e-3 (relative L2);
 * surrogate: lifetimes within 3e-3 (median) and 1.5e-2 (p99), evaporated
   mass within 2e-3 of the initial liquid mass at every sample, mass-source
   field within 1e-2 (relative L2), more than 95 % of the droplet-substeps
   served by the network."""
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from pelemp_amd import _capi as C

    if C.device_count() < 1:
        pytest.skip("no HIP device")
    return C


@pytest.mark.parametrize("recipe", ["config4", "config2"])
def test_errors_over_droplet_lifetimes(recipe):
    C = _lib()
    import bench
    from pelemp_amd import problem as P

    prob = P.config4(ncell=12) if recipe == "config4" else P.config2(ncell=12)
    steps = 8000 if recipe == "config4" else 16000
    every = 10 if recipe == "config4" else 20
    print(f"lifetimes: {recipe}, {prob.n} droplets, up to {steps} steps", flush=True)
    r = bench.lifetime_study(prob, C.Stream(), max_steps=steps, every=every, field_steps=(steps // 8,))
    assert r["fp64"]["evaporated_frac_end"] > 0.9
    f32, sur = r["fp32"], r["surrogate"]
    assert f32["lifetime_droplets_compared"] > 0.9 * prob.n
    assert f32["lifetime_rel_err"]["median"] < 1e-5 and f32["lifetime_rel_err"]["p99"] < 1e-4
    assert f32["evaporated_frac_max_abs_err"] < 1e-5
    (fe,) = f32["source_field_err"].values()
    assert fe["l2"]["mass"] < 1e-3
    assert sur["surrogate_share"] > 0.95
    assert sur["lifetime_rel_err"]["median"] < 3e-3 and sur["lifetime_rel_err"]["p99"] < 1.5e-2
    assert sur["evaporated_frac_max_abs_err"] < 2e-3
    (fe,) = sur["source_field_err"].values()
    assert fe["l2"]["mass"] < 1e-2
