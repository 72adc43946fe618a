// particle_io.h -- AMReX's particle checkpoint layout on plain host arrays.
//
// SprayParticleIO (SprayIO.cpp:6-34) writes the droplets through AMReX's
// ParticleContainer::Checkpoint(dir, "particles", ...) and a restart reads
// them back through ParticleContainer::Restart(dir, "particles")
// (SpraySetup.cpp:320-325).  AMReX is a dependency that the reference does
// not vendor; the layout below is the one of its binary particle writer
// (AMReX_WriteBinaryParticleData.H, AMReX_ParticleIO.H):
//
//   <dir>/<name>/Header            text:
//       Version_Two_Dot_One_double   (checkpoints; "_single" for float reals)
//       3                            space dimension
//       nr, then nr names            real components beyond the position
//       ni, then ni names            int components beyond id and cpu
//       1                            the legacy "is checkpoint" flag
//       nparticles
//       next id                      ParticleType::NextID() on restart
//       finest level
//       the grid count of each level
//       per level, per grid: "which count where" (DATA file number, particle
//       count, byte offset of the grid's block in that file)
//   <dir>/<name>/Level_L/Particle_H  the level's BoxArray (BoxArray::writeOn)
//   <dir>/<name>/Level_L/DATA_nnnnn  per grid with particles: count x (2+ni)
//       int32 (the 64-bit idcpu word split high/low for Version_Two_Dot_One,
//       plain id and cpu for Version_Two_Dot_Zero), then count x (3+nr)
//       reals (position, then the components in order), native byte order.
//
// idcpu packs the id in sign-magnitude form (bit 63 set for id >= 0, |id| in
// bits 24-62) and the creating rank in bits 0-23.  The context holds one
// 64-bit id per droplet; a droplet made on rank r > 0 is held as
// id + r * 2^39 (ids are below 2^39 in AMReX), so (id, cpu) pairs stay
// unique and are written back unchanged.
#pragma once

#include <sys/stat.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "pelemp_spray.h"

namespace pelemp {
namespace io {

constexpr int kIdBits = 39;
constexpr int64_t kIdMask = (int64_t(1) << kIdBits) - 1;
constexpr int kCpuMax = (1 << 24) - 1;

inline int64_t amrex_id(int64_t held) { return held & kIdMask; }
inline int amrex_cpu(int64_t held) { return static_cast<int>(held >> kIdBits); }
inline int64_t held_id(int64_t id, int cpu) { return (static_cast<int64_t>(cpu) << kIdBits) | id; }

inline void pack_idcpu(int64_t id, int cpu, int32_t out[2]) {
  const uint64_t w = (uint64_t(1) << 63) | (static_cast<uint64_t>(id) << 24) | static_cast<uint64_t>(cpu);
  const uint32_t hi = static_cast<uint32_t>(w >> 32), lo = static_cast<uint32_t>(w & 0xFFFFFFFFu);
  std::memcpy(&out[0], &hi, 4);
  std::memcpy(&out[1], &lo, 4);
}

// false for a particle that AMReX marks invalid (negative id)
inline bool unpack_idcpu(const int32_t in[2], int64_t* id, int* cpu) {
  uint32_t hi, lo;
  std::memcpy(&hi, &in[0], 4);
  std::memcpy(&lo, &in[1], 4);
  const uint64_t w = (static_cast<uint64_t>(hi) << 32) | lo;
  *cpu = static_cast<int>(w & 0xFFFFFFu);
  *id = static_cast<int64_t>((w >> 24) & 0x7FFFFFFFFFull);
  return (w >> 63) != 0;
}

struct Box {
  int lo[3], hi[3];
};

inline std::string grid_file(const std::string& level_dir, int which) {
  char buf[32];
  std::snprintf(buf, sizeof(buf), "/DATA_%05d", which);
  return level_dir + buf;
}

// Writes level 0 of a checkpoint.  Particles are assigned to the first box
// that holds their cell; err names a particle that no box holds.
inline int write_checkpoint(const std::string& dir, const std::string& name, int64_t n, const double* pos,
                            const double* rd, const int64_t* ids, int nreal, const std::vector<std::string>& names,
                            int64_t next_id, const std::vector<Box>& boxes, const double plo[3],
                            const double dx[3], std::string& err) {
  if (n < 0 || nreal < 0 || static_cast<int>(names.size()) != nreal || boxes.empty() || (n > 0 && (!pos || !ids)) ||
      (n > 0 && nreal > 0 && !rd)) {
    err = "particle checkpoint: bad arguments";
    return PELEMP_E_ARG;
  }
  for (const auto& nm : names)
    if (nm.empty() || nm.find_first_of(" \t\n") != std::string::npos) {
      err = "particle checkpoint: component names must be single words";
      return PELEMP_E_ARG;
    }
  const int nb = static_cast<int>(boxes.size());
  std::vector<int> grid(static_cast<size_t>(n));
  std::vector<int64_t> count(nb, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (ids[i] < 0 || amrex_cpu(ids[i]) > kCpuMax) {
      err = "particle checkpoint: id " + std::to_string(ids[i]) + " has no AMReX (id, cpu) form";
      return PELEMP_E_ARG;
    }
    int g = -1;
    if (nb == 1) {
      g = 0;
    } else {
      int cell[3];
      for (int d = 0; d < 3; ++d)
        cell[d] = static_cast<int>(std::floor((pos[3 * i + d] - plo[d]) / dx[d]));
      for (int b = 0; b < nb && g < 0; ++b) {
        bool in = true;
        for (int d = 0; d < 3; ++d) in = in && cell[d] >= boxes[b].lo[d] && cell[d] <= boxes[b].hi[d];
        if (in) g = b;
      }
      if (g < 0) {
        err = "particle checkpoint: particle " + std::to_string(ids[i]) + " lies in no box of the box array";
        return PELEMP_E_ARG;
      }
    }
    grid[i] = g;
    ++count[g];
  }
  const std::string pdir = dir + "/" + name;
  mkdir(dir.c_str(), 0755);
  mkdir(pdir.c_str(), 0755);
  std::vector<int64_t> where(nb, 0);
  if (n > 0) {
    const std::string ldir = pdir + "/Level_0";
    mkdir(ldir.c_str(), 0755);
    {
      std::ofstream ph(ldir + "/Particle_H", std::ios::out | std::ios::trunc);
      ph << '(' << nb << ' ' << 0 << '\n';
      for (const auto& b : boxes)
        ph << "((" << b.lo[0] << ',' << b.lo[1] << ',' << b.lo[2] << ") (" << b.hi[0] << ',' << b.hi[1] << ','
           << b.hi[2] << ") (0,0,0))\n";
      ph << ')' << '\n';
      ph.close();
      if (!ph.good()) {
        err = "particle checkpoint: cannot write " + ldir + "/Particle_H";
        return PELEMP_E_ARG;
      }
    }
    std::ofstream df(grid_file(ldir, 0), std::ios::out | std::ios::trunc | std::ios::binary);
    const int rc = 3 + nreal;
    // particles grouped by grid, in their given order within a grid
    std::vector<int64_t> start(nb + 1, 0), order(static_cast<size_t>(n));
    for (int g = 0; g < nb; ++g) start[g + 1] = start[g] + count[g];
    {
      std::vector<int64_t> fill(start.begin(), start.end() - 1);
      for (int64_t i = 0; i < n; ++i) order[fill[grid[i]]++] = i;
    }
    std::vector<int32_t> ib;
    std::vector<double> rb;
    for (int g = 0; g < nb; ++g) {
      where[g] = static_cast<int64_t>(df.tellp());
      if (count[g] == 0) continue;
      ib.assign(2 * count[g], 0);
      rb.assign(rc * count[g], 0.0);
      for (int64_t k = 0; k < count[g]; ++k) {
        const int64_t i = order[start[g] + k];
        pack_idcpu(amrex_id(ids[i]), amrex_cpu(ids[i]), &ib[2 * k]);
        for (int d = 0; d < 3; ++d) rb[rc * k + d] = pos[3 * i + d];
        for (int q = 0; q < nreal; ++q) rb[rc * k + 3 + q] = rd[static_cast<int64_t>(nreal) * i + q];
      }
      df.write(reinterpret_cast<const char*>(ib.data()), static_cast<std::streamsize>(4 * ib.size()));
      df.write(reinterpret_cast<const char*>(rb.data()), static_cast<std::streamsize>(8 * rb.size()));
    }
    df.close();
    if (!df.good()) {
      err = "particle checkpoint: cannot write " + grid_file(ldir, 0);
      return PELEMP_E_ARG;
    }
  }
  std::ofstream hd(pdir + "/Header", std::ios::out | std::ios::trunc);
  hd << "Version_Two_Dot_One_double\n" << 3 << '\n' << nreal << '\n';
  for (const auto& nm : names) hd << nm << '\n';
  hd << 0 << '\n' << 1 << '\n' << n << '\n' << next_id << '\n' << 0 << '\n' << nb << '\n';
  for (int g = 0; g < nb; ++g) hd << 0 << ' ' << count[g] << ' ' << where[g] << '\n';
  hd.close();
  if (!hd.good()) {
    err = "particle checkpoint: cannot write " + pdir + "/Header";
    return PELEMP_E_ARG;
  }
  return PELEMP_OK;
}

struct HeaderInfo {
  bool ids_packed = true, single = false;
  int nreal = 0, nint = 0, finest = 0;
  int64_t nparticles = 0, next_id = 0;
  std::vector<std::string> names;
  std::vector<int> ngrids;
  // per level, per grid
  std::vector<std::vector<int>> which;
  std::vector<std::vector<int64_t>> count, where;
};

inline int read_header(const std::string& pdir, HeaderInfo& h, std::string& err) {
  std::ifstream hd(pdir + "/Header");
  if (!hd.good()) {
    err = "not a particle checkpoint (no Header): " + pdir;
    return PELEMP_E_ARG;
  }
  std::string version;
  hd >> version;
  // Version_Two_Dot_Zero: plain id and cpu; Version_Two_Dot_One: the idcpu word
  if (version.rfind("Version_Two_Dot_One", 0) == 0) {
    h.ids_packed = true;
  } else if (version.rfind("Version_Two_Dot_Zero", 0) == 0) {
    h.ids_packed = false;
  } else {
    err = "particle checkpoint: unsupported version string " + version;
    return PELEMP_E_ARG;
  }
  if (version.size() > 7 && version.compare(version.size() - 7, 7, "_single") == 0) {
    h.single = true;
  } else if (version.size() > 7 && version.compare(version.size() - 7, 7, "_double") == 0) {
    h.single = false;
  } else {
    err = "particle checkpoint: bad version string " + version;
    return PELEMP_E_ARG;
  }
  int dm = 0, is_ckpt = 0;
  hd >> dm >> h.nreal;
  if (!hd.good() || dm != 3 || h.nreal < 0 || h.nreal > 4096) {
    err = "particle checkpoint: bad dimension or component count in " + pdir + "/Header";
    return PELEMP_E_ARG;
  }
  h.names.resize(h.nreal);
  for (auto& nm : h.names) hd >> nm;
  hd >> h.nint;
  if (!hd.good() || h.nint < 0 || h.nint > 4096) {
    err = "particle checkpoint: bad int component count";
    return PELEMP_E_ARG;
  }
  for (int i = 0; i < h.nint; ++i) {
    std::string s;
    hd >> s;
  }
  hd >> is_ckpt >> h.nparticles >> h.next_id >> h.finest;
  if (!hd.good() || h.nparticles < 0 || h.finest < 0 || h.finest > 64) {
    err = "particle checkpoint: truncated Header";
    return PELEMP_E_ARG;
  }
  h.ngrids.resize(h.finest + 1);
  for (auto& g : h.ngrids) hd >> g;
  h.which.resize(h.finest + 1);
  h.count.resize(h.finest + 1);
  h.where.resize(h.finest + 1);
  for (int lev = 0; lev <= h.finest; ++lev) {
    const int ng = h.ngrids[lev];
    if (!hd.good() || ng < 0) {
      err = "particle checkpoint: bad grid count";
      return PELEMP_E_ARG;
    }
    h.which[lev].resize(ng);
    h.count[lev].resize(ng);
    h.where[lev].resize(ng);
    for (int g = 0; g < ng; ++g) hd >> h.which[lev][g] >> h.count[lev][g] >> h.where[lev][g];
  }
  if (hd.fail()) {
    err = "particle checkpoint: truncated grid list in " + pdir + "/Header";
    return PELEMP_E_ARG;
  }
  return PELEMP_OK;
}

inline int64_t level_count(const HeaderInfo& h, int level) {
  int64_t n = 0;
  for (int64_t c : h.count[level]) n += c;
  return n;
}

// Reads one level's particles; ids in the held form (id + cpu * 2^39).
// Particles with a negative id (invalid in AMReX) are skipped.
inline int read_level(const std::string& pdir, const HeaderInfo& h, int level, std::vector<double>& pos,
                      std::vector<double>& rd, std::vector<int64_t>& ids, std::string& err) {
  if (level < 0 || level > h.finest) {
    err = "particle checkpoint: level " + std::to_string(level) + " is not in the checkpoint";
    return PELEMP_E_ARG;
  }
  const int ic = 2 + h.nint, rc = 3 + h.nreal;
  const size_t rsz = h.single ? 4 : 8;
  const std::string ldir = pdir + "/Level_" + std::to_string(level);
  pos.clear();
  rd.clear();
  ids.clear();
  std::vector<int32_t> ib;
  std::vector<char> rb;
  for (size_t g = 0; g < h.count[level].size(); ++g) {
    const int64_t cnt = h.count[level][g];
    if (cnt == 0) continue;
    if (cnt < 0 || h.where[level][g] < 0) {
      err = "particle checkpoint: bad grid entry";
      return PELEMP_E_ARG;
    }
    const std::string fn = grid_file(ldir, h.which[level][g]);
    std::ifstream df(fn, std::ios::binary);
    df.seekg(h.where[level][g]);
    ib.resize(static_cast<size_t>(ic) * cnt);
    rb.resize(rsz * rc * cnt);
    df.read(reinterpret_cast<char*>(ib.data()), static_cast<std::streamsize>(4 * ib.size()));
    df.read(rb.data(), static_cast<std::streamsize>(rb.size()));
    if (!df.good()) {
      err = "particle checkpoint: truncated or missing " + fn;
      return PELEMP_E_ARG;
    }
    for (int64_t k = 0; k < cnt; ++k) {
      int64_t id;
      int cpu;
      const int32_t* ip = &ib[static_cast<size_t>(ic) * k];
      if (h.ids_packed) {
        if (!unpack_idcpu(ip, &id, &cpu)) continue;
      } else {
        if (ip[0] < 0) continue;
        id = ip[0];
        cpu = ip[1];
      }
      if (id > kIdMask || cpu < 0 || cpu > kCpuMax) {
        err = "particle checkpoint: particle id out of range in " + fn;
        return PELEMP_E_ARG;
      }
      ids.push_back(held_id(id, cpu));
      const char* r = &rb[rsz * rc * k];
      for (int q = 0; q < rc; ++q) {
        double v;
        if (h.single) {
          float f;
          std::memcpy(&f, r + 4 * q, 4);
          v = f;
        } else {
          std::memcpy(&v, r + 8 * q, 8);
        }
        (q < 3 ? pos : rd).push_back(v);
      }
    }
  }
  return PELEMP_OK;
}

}  // namespace io
}  // namespace pelemp
