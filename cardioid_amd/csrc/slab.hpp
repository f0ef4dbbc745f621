// slab.hpp -- host preparation of one tissue slab for the engine.
//
// A slab is one anatomy (the cells of a rank in the reference): its cells in
// box order (x-major, z fastest), so that a cell's stencil neighbours are
// near it in memory, padded to whole 64-cell tiles; per-cell metadata; FGR
// weights; rest states; the initial box of Vm.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/cardioid_hip.h"
#include "fgr_weights.hpp"
#include "tt06_f64.hpp"
#include "tt06_model.hpp"

namespace cdh {

struct SlabInput {
   int nx = 0, ny = 0, nz = 0;
   double dx = 0, dy = 0, dz = 0;
   double diffusionScale = 0;
   int64_t ncell = 0;
   const int64_t *gid = nullptr;       // ncell grid indices x + nx*(y + ny*z)
   const uint8_t *celltype = nullptr;  // ncell, index into celltypes
   const double *sigma = nullptr;      // 6*ncell
   int ncelltype = 0;
   const CellType *celltypes = nullptr;
   const double *scales = nullptr;     // 3*ncell (s_kr, s_cal, s_na) or null
   double threshold = -40.0;
   int gate_tail_rule = 1;
   int nstim = 0;
   const cdh_box_stim *stim = nullptr;
   int check_ranges = 1;               // Simulate::checkRanges with [vm_min, vm_max]
   double vm_min = -150.0, vm_max = 100.0;
   int setup_on_host = 0;              // 0: the weights are built on the device
   // remote cells (the reference's nRemote): tissue the slab's stencil reads
   // but does not update; their Vm comes from the slab that owns them
   int64_t nremote = 0;
   const int64_t *remote_gid = nullptr;
   const double *remote_sigma = nullptr;  // 6*nremote
};

struct SlabHost {
   BoxGeom geom{};
   int64_t n = 0;                      // tissue cells
   int64_t ntile = 0;                  // 64-cell tiles
   int64_t S = 0;                      // ntile * 64
   int64_t nbox = 0;                   // box points
   std::vector<int64_t> perm;          // slab order -> input order
   std::vector<uint32_t> bi;           // S: box index (padding cells: cell 0's)
   std::vector<uint32_t> meta;         // S: meta_pack(pid, gate flags, stimulus mask)
   std::vector<ParamSet> psets;        // parameter sets, indexed by pid
   std::vector<double> state;          // [19][S]: rest states, dVK_i = K_i/c9 + Vm0
   std::vector<double> A0;             // S
   std::vector<float> Aw;              // [18][S]
   // compact weights of the mixed form, in units of 2^-wexp (so that the
   // largest fits FP16's range): per 64-cell tile a reference weight per
   // neighbour, per cell the difference from it in FP16; two neighbours per
   // 32-bit word ([9][S], low half = even neighbour)
   std::vector<float> Aref;            // [ntile][18]
   std::vector<uint32_t> Aw16;         // [9][S]
   float wdelta_scale = 1.0f;          // 2^-wexp, the unit of Aref and Aw16
   // device set-up: the weights above are left empty and the engine builds
   // them from the cells' conductivities ([n][6], slab order)
   bool device_weights = false;
   std::vector<double> sig_cells;
   int gnx = 0, gny = 0;
   std::vector<int64_t> remote_gid;
   std::vector<uint32_t> remote_bi;     // box index within the slab's box
   std::vector<double> remote_sig;      // [nremote][6]
   std::vector<double> box0;           // nbox: rest Vm at tissue points, 0 elsewhere
   int64_t off[19] = {0};              // NodeLocation offsets in the box
   std::vector<cdh_box_stim> stim;
   std::vector<std::vector<int>> stimList;   // slab-order cells per stimulus
   double diffusionScale = 0, threshold = -40.0;
   int check_ranges = 1;
   double vm_min = -150.0, vm_max = 100.0;
};

// Validates the input and builds the slab (elec/FGRDiffusionOMP.cc:14-73
// for the box and weights, elec/ReactionManager.cc:218-261 for the gate
// groups, elec/BoxStimulus.cc:18-34 for stimulus membership).
SlabHost prepare_slab(const SlabInput &in, double c9);

// tiles and box points a slab needs, without building it
void slab_extent(const SlabInput &in, int64_t *ntile, int64_t *nbox);

}  // namespace cdh
