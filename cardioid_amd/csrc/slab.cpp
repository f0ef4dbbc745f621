// slab.cpp -- host preparation of one tissue slab (see slab.hpp).
#include "slab.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>

#include "engine.hpp"
#include "fgr_stencil.hpp"
#include "hip_util.hpp"
#include "setup_device.hpp"
#include "setup.hpp"

namespace cdh {

namespace {

void check_input(const SlabInput &in)
{
   CDH_REQUIRE(in.ncell >= 1 && in.ncell < INT32_MAX, "slab: ncell must be 1..2^31-1");
   CDH_REQUIRE(in.gid && in.celltype && in.sigma, "slab: null cell arrays");
   CDH_REQUIRE(in.nx > 0 && in.ny > 0 && in.nz > 0, "slab: grid dimensions must be positive");
   CDH_REQUIRE(in.dx > 0 && in.dy > 0 && in.dz > 0, "slab: grid spacing must be positive");
   CDH_REQUIRE(in.ncelltype >= 1 && in.ncelltype <= kMaxCellTypes && in.celltypes,
               "slab: 1..8 cell types required");
   CDH_REQUIRE(in.nstim >= 0 && in.nstim <= kSlabStim && (in.nstim == 0 || in.stim),
               "slab: at most 8 stimuli");
   const int64_t npts = (int64_t)in.nx * in.ny * in.nz;
   for (int64_t c = 0; c < in.ncell; c++) {
      CDH_REQUIRE(in.gid[c] >= 0 && in.gid[c] < npts, "slab: gid outside the grid");
      CDH_REQUIRE(in.celltype[c] < in.ncelltype, "slab: cell type index out of range");
   }
   for (int q = 0; q < in.nstim; q++)
      CDH_REQUIRE(in.stim[q].period > 0 && in.stim[q].duration > 0, "slab: bad stimulus pulse");
   CDH_REQUIRE(in.nremote >= 0 && (in.nremote == 0 || (in.remote_gid && in.remote_sigma)),
               "slab: null remote cell arrays");
   for (int64_t c = 0; c < in.nremote; c++)
      CDH_REQUIRE(in.remote_gid[c] >= 0 && in.remote_gid[c] < npts,
                  "slab: remote gid outside the grid");
}

// Per tile and neighbour a reference weight, per cell its difference from
// that reference in FP16 after scaling by 2^wexp (chosen so that the slab's
// largest weight times 2^wexp is at most 2^14: differences stay below FP16's
// largest finite value, and a difference 2^-28 of the largest weight is
// still a normal FP16 number).  The reference is the mean of the tile's
// nonzero weights, rounded to a value whose scaled form FP16 holds exactly:
// a zero weight (a neighbour outside the tissue, a dropped transverse term)
// is then stored as the exact negative of the reference and rebuilt as an
// exact zero by ref + diff*2^-wexp in one fused multiply-add.
void compact_weights(SlabHost &h)
{
   const int64_t S = h.S;
   double wmax = 0.0;
   for (float w : h.Aw) wmax = std::max(wmax, (double)std::fabs(w));
   const int wexp = compact_weight_exponent((float)wmax);   // wmax is a float
   const double up = std::ldexp(1.0, wexp);
   h.wdelta_scale = (float)std::ldexp(1.0, -wexp);
   h.Aref.assign((size_t)h.ntile * 18, 0.0f);
   h.Aw16.assign((size_t)9 * S, 0u);
#pragma omp parallel for schedule(static)
   for (int64_t t = 0; t < h.ntile; t++) {
      const int64_t c0 = t * kBlk, c1 = std::min<int64_t>(c0 + kBlk, h.n);
      for (int k = 0; k < 18; k++) {
         double sum = 0.0;
         int64_t nz = 0;
         for (int64_t c = c0; c < c1; c++)
            if (h.Aw[k * S + c] != 0.0f) {
               sum += h.Aw[k * S + c];
               ++nz;
            }
         const float ref = (float)fgr::half_rne(nz ? sum / (double)nz * up : 0.0);   // units of 2^-wexp
         h.Aref[t * 18 + k] = ref;
         for (int64_t c = c0; c < c0 + kBlk; c++) {
            const double d = (double)h.Aw[k * S + c] * up - (double)ref;
            const uint16_t bits = fgr::half_bits(fgr::half_rne(d));
            h.Aw16[(k / 2) * S + c] |= (uint32_t)bits << (16 * (k & 1));
         }
      }
   }
}

}  // namespace

void slab_extent(const SlabInput &in, int64_t *ntile, int64_t *nbox)
{
   check_input(in);
   const BoxGeom g = bounding_box(in.nx, in.ny, in.nz, in.gid, in.ncell, in.dx, in.dy, in.dz);
   *ntile = (in.ncell + kBlk - 1) / kBlk;
   *nbox = (int64_t)g.bx * g.by * g.bz;
   CDH_REQUIRE(*nbox < (int64_t(1) << 28), "slab: box above 2^28 points");
   CDH_REQUIRE((int64_t)g.by * g.bz + g.bz + 1 < kBoxShift, "slab: box plane above 2^22 points");
}

SlabHost prepare_slab(const SlabInput &in, double c9)
{
   check_input(in);
   SlabHost h;
   const int64_t n = in.ncell;
   h.n = n;
   h.ntile = (n + kBlk - 1) / kBlk;
   h.S = h.ntile * kBlk;
   h.diffusionScale = in.diffusionScale;
   h.threshold = in.threshold;
   h.check_ranges = in.check_ranges;
   h.vm_min = in.vm_min;
   h.vm_max = in.vm_max;
   const int64_t S = h.S;

   // box and slab order
   h.geom = bounding_box(in.nx, in.ny, in.nz, in.gid, n, in.dx, in.dy, in.dz);
   const BoxGeom &g = h.geom;
   h.nbox = (int64_t)g.bx * g.by * g.bz;
   CDH_REQUIRE(h.nbox < (int64_t(1) << 28), "slab: box above 2^28 points");
   CDH_REQUIRE((int64_t)g.by * g.bz + g.bz + 1 < kBoxShift, "slab: box plane above 2^22 points");
   std::vector<int64_t> bidxIn(n);
   for (int64_t c = 0; c < n; c++) bidxIn[c] = box_index(g, in.nx, in.ny, in.gid[c]);
   h.perm.resize(n);
   std::iota(h.perm.begin(), h.perm.end(), 0);
   std::sort(h.perm.begin(), h.perm.end(),
             [&](int64_t a, int64_t b) { return bidxIn[a] < bidxIn[b]; });
   for (int64_t k = 1; k < n; k++)
      CDH_REQUIRE(bidxIn[h.perm[k]] != bidxIn[h.perm[k - 1]], "slab: duplicate gid");

   // gate flags from each cell's position in its reaction group (cell type)
   // in the input order; parameter sets per (cell type, scales)
   std::vector<int64_t> groupCount(in.ncelltype, 0), groupPos(n);
   for (int64_t c = 0; c < n; c++) groupPos[c] = groupCount[in.celltype[c]]++;
   std::map<std::tuple<int, double, double, double>, int> psetId;
   std::vector<uint8_t> pidIn(n);
   for (int64_t c = 0; c < n; c++) {
      const double skr = in.scales ? in.scales[3 * c] : 1.0;
      const double scal = in.scales ? in.scales[3 * c + 1] : 1.0;
      const double sna = in.scales ? in.scales[3 * c + 2] : 1.0;
      auto key = std::make_tuple((int)in.celltype[c], skr, scal, sna);
      auto it = psetId.find(key);
      if (it == psetId.end()) {
         CDH_REQUIRE(h.psets.size() < (size_t)kMaxParamSets,
                     "slab: more than 256 (cell type, scale) combinations");
         it = psetId.emplace(key, (int)h.psets.size()).first;
         h.psets.push_back(make_param_set(in.celltypes[in.celltype[c]], skr, scal, sna));
      }
      pidIn[c] = (uint8_t)it->second;
   }

   // per-cell arrays in slab order; padding cells alias cell 0's box point,
   // keep their gates fixed and never store Vm
   h.bi.assign(S, 0);
   h.meta.assign(S, 0);
   h.state.assign((size_t)N_STATE * S, 0.0);
   h.box0.assign(h.nbox, 0.0);
   // host set-up only: tissue and conductivity over the box
   const bool host = in.setup_on_host != 0;
   std::vector<uint8_t> tissue(host ? h.nbox : 0, 0);
   std::vector<double> sig(host ? 6 * h.nbox : 0, 0.0);
   std::vector<int64_t> bidx(n);
   if (host) {
      h.A0.assign(S, 0.0);
      h.Aw.assign((size_t)18 * S, 0.0f);
   } else {
      h.device_weights = true;
      h.sig_cells.resize((size_t)6 * n);
   }
   std::vector<uint8_t> stimMask(S, 0);
   h.stim.assign(in.stim, in.stim + in.nstim);
   h.stimList.assign(in.nstim, {});
#pragma omp parallel for schedule(static)
   for (int64_t k = 0; k < n; k++) {
      const int64_t c = h.perm[k];
      bidx[k] = bidxIn[c];
      h.bi[k] = (uint32_t)bidxIn[c];
      if (host) {
         tissue[bidxIn[c]] = 1;
         std::memcpy(&sig[6 * bidxIn[c]], in.sigma + 6 * c, 6 * sizeof(double));
      } else {
         std::memcpy(&h.sig_cells[6 * k], in.sigma + 6 * c, 6 * sizeof(double));
      }
      const CellType &ct = in.celltypes[in.celltype[c]];
      for (int q = 0; q < N_STATE; q++) h.state[q * S + k] = ct.state0[q];
      h.state[S_DVK_I * S + k] = ct.state0[S_DVK_I] / c9 + ct.Vm0;
      h.box0[bidxIn[c]] = ct.Vm0;
      // strict box in grid tuples (elec/BoxStimulus.cc:18-34)
      const int64_t gg = in.gid[c];
      const int64_t x = gg % in.nx, y = (gg / in.nx) % in.ny, z = gg / ((int64_t)in.nx * in.ny);
      for (int q = 0; q < in.nstim; q++) {
         const cdh_box_stim &b = in.stim[q];
         if (x > b.xMin && x < b.xMax && y > b.yMin && y < b.yMax && z > b.zMin && z < b.zMax)
            stimMask[k] |= (uint8_t)(1u << q);
      }
      const uint8_t fl = in.gate_tail_rule
                             ? gate_flags(groupPos[c], groupCount[in.celltype[c]])
                             : (uint8_t)0;
      h.meta[k] = meta_pack(pidIn[c], fl, stimMask[k]);
   }
   for (int64_t k = 0; k < n; k++)
      for (int q = 0; q < in.nstim; q++)
         if ((stimMask[k] >> q) & 1u) h.stimList[q].push_back((int)k);
   const CellType &ct0 = in.celltypes[in.celltype[h.perm[0]]];
   for (int64_t k = n; k < S; k++) {
      h.bi[k] = h.bi[0];
      h.meta[k] = meta_pack(pidIn[h.perm[0]], kFlagGatesFixed, 0) | kMetaPad;
      for (int q = 0; q < N_STATE; q++) h.state[q * S + k] = ct0.state0[q];
      h.state[S_DVK_I * S + k] = ct0.state0[S_DVK_I] / c9 + ct0.Vm0;
   }
   // remote cells: inside the padded box, not local, each once
   h.gnx = in.nx;
   h.gny = in.ny;
   if (in.nremote > 0) {
      std::vector<uint8_t> taken(h.nbox, 0);
      for (int64_t k = 0; k < n; k++) taken[h.bi[k]] = 1;
      h.remote_gid.assign(in.remote_gid, in.remote_gid + in.nremote);
      h.remote_bi.resize(in.nremote);
      for (int64_t c = 0; c < in.nremote; c++) {
         const int64_t rb = box_index(g, in.nx, in.ny, in.remote_gid[c]);
         CDH_REQUIRE(rb >= 0, "slab: a remote cell lies outside the padded box");
         CDH_REQUIRE(!taken[rb], "slab: a remote cell is local or listed twice");
         taken[rb] = 1;
         h.remote_bi[c] = (uint32_t)rb;
         if (host) {
            tissue[rb] = 1;
            std::memcpy(&sig[6 * rb], in.remote_sigma + 6 * c, 6 * sizeof(double));
         }
      }
      if (!host) h.remote_sig.assign(in.remote_sigma, in.remote_sigma + 6 * in.nremote);
   }
   fgr_offsets(g.bx, g.by, g.bz, h.off);
   if (!host) return h;
   fgr_weights(g, tissue.data(), sig.data(), n, bidx.data(), h.A0.data(), h.Aw.data(), S);
   for (int64_t k = n; k < S; k++) {
      h.A0[k] = h.A0[0];
      for (int q = 0; q < 18; q++) h.Aw[q * S + k] = h.Aw[q * S];
   }
   compact_weights(h);
   return h;
}

}  // namespace cdh
