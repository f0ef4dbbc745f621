// plugins.hip -- the drop-in objects behind the C ABI: the TT06Dev reaction,
// the FGR diffusion, the loop helpers and the activation sensor.
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cardioid_hip.h"
#include "fgr_weights.hpp"
#include "hip_util.hpp"
#include "setup_device.hpp"
#include "kernels.hpp"
#include "row_shift.hpp"
#include "setup.hpp"
#include "tt06_f64.hpp"

namespace cdh {

namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string &msg) { g_last_error = msg; }

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Reaction
// ---------------------------------------------------------------------------
struct TT06Reaction {
   int n = 0;
   int64_t stride = 0;
   bool exact = false;              // configuration A: tolerance 0, the exact functions
   TT06Const K;
   DevBuf<TT06Const> Kd;            // the same constants on the device
   DevBuf<TT06ExactConst> Ed;       // configuration A's constants
   CellType ct;
   DevBuf<double> coef;
   DevBuf<ParamSet> params;
   DevBuf<uint8_t> flags;
   DevBuf<double> state;
   DevBuf<uint8_t> flagsAllFixed;   // gates skipped (updateNonGate-only launches)
   int64_t pipe_blocks = 0;         // resident workgroups of k_tt06_calc_pipe_f64
   DevBuf<double> sink;             // its stores of lanes past n (20 x 64 doubles)
};

}  // namespace cdh

using namespace cdh;

struct cdh_tt06 : TT06Reaction {};
struct cdh_fgr {
   int64_t nLocal = 0, nRemote = 0;
   int64_t stride = 0;
   BoxGeom geom;
   StencilArgs args;
   DevBuf<uint32_t> bi;       // local then remote
   DevBuf<double> box, A0;
   DevBuf<float> Aw;
   double scale = 0.0;
};
struct cdh_sensor {
   int64_t n = 0;
   double threshold = -40.0;
   int evaluated = 0;
   DevBuf<double> tact0, tact, trec;
   DevBuf<uint8_t> active, activated;
   int apd = 0;
   DevBuf<uint8_t> apd_phase;
   DevBuf<double> vmax, t90, vrest;
   DevBuf<uint32_t> counts;   // recorded tact0, tact, trec (and one unused word)
};

namespace {

hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

ReactionArgs reaction_args(cdh_tt06 *r, double dt, const int *index, int n, const double *Vm,
                           double *dVm, const uint8_t *flags)
{
   ReactionArgs a;
   a.K = r->Kd.p;
   a.coef = r->coef.p;
   a.params = r->params.p;
   a.pid = nullptr;
   a.flags = flags;
   a.state = r->state.p;
   a.n = n;
   a.index = index;
   a.Vm = Vm;
   a.dVm = dVm;
   a.dt = dt;
   a.gate_mask = 0x1FFFu;
   a.E = r->exact ? r->Ed.p : nullptr;
   a.sink = r->sink.p;
   return a;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// box[bi[c]] = src[c] for n cells: four cells per thread where both arrays
// allow 16-byte loads
void scatter(int64_t n, const uint32_t *bi, const double *src, double *box, hipStream_t s)
{
   if (aligned16(bi) && aligned16(src))
      hipLaunchKernelGGL(k_scatter4_f64, dim3(grid_for((n + 3) / 4, kBlock)), dim3(kBlock), 0, s,
                         (int)n, bi, src, box);
   else
      hipLaunchKernelGGL(k_scatter_f64, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, (int)n, bi,
                         src, box);
   CDH_LAUNCH_CHECK();
}

// The reaction over the object's n cells: configuration A's kernel, or the
// pipelined kernel of configuration B on its resident grid.
void launch_calc(cdh_tt06 *r, const ReactionArgs &a, hipStream_t s)
{
   if (r->exact) {
      hipLaunchKernelGGL(k_tt06_calc_exact_f64, dim3(grid_for(a.n, kBlock)), dim3(kBlock), 0, s,
                         a);
   } else {
      const int64_t tiles = ((int64_t)a.n + kBlk - 1) / kBlk;
      const int64_t blocks = std::min<int64_t>((tiles + kBlock / 64 - 1) / (kBlock / 64),
                                               r->pipe_blocks);
      hipLaunchKernelGGL(k_tt06_calc_pipe_f64, dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
   }
   CDH_LAUNCH_CHECK();
}

const char *kFieldName[N_STATE] = {"Ca_i", "dVK_i", "Na_i", "Ca_ss", "Ca_SR", "R_prime",
                                   "fCass_gate", "m_gate", "h_gate", "j_gate", "Xr1_gate",
                                   "Xr2_gate", "Xs_gate", "r_gate", "d_gate", "f_gate",
                                   "f2_gate", "jL_gate", "s_gate"};
// units as TT06Dev_Reaction::getHandleMap lists them (:533-551)
const char *kFieldUnit[N_STATE] = {"mM", "mV", "mM", "mM", "mM", "1", "1", "1", "1", "1",
                                   "1", "1", "1", "1", "1", "1", "1", "mM", "1"};

}  // namespace

extern "C" {

const char *cdh_last_error(void) { return g_last_error.c_str(); }
int cdh_version(void) { return 1; }

int cdh_celltype_standard(const char *name, cdh_celltype *out)
{
   return guard([&] {
      CDH_REQUIRE(name && out, "cdh_celltype_standard: null argument");
      CellType ct;
      if (!tt06_standard_celltype(name, &ct))
         throw Error(kErrArg, std::string("unknown cell type ") + name);
      static_assert(sizeof(cdh_celltype) == sizeof(CellType), "cell type layout");
      std::memcpy(out, &ct, sizeof(ct));
   });
}

int cdh_tt06_create(const cdh_tt06_params *p, int numPoints, cdh_tt06 **out)
{
   return guard([&] {
      CDH_REQUIRE(p && out && numPoints >= 0, "cdh_tt06_create: bad arguments");
      // tolerance 0 (the factory's default) keeps the exact functions and
      // the generic updates whatever fastGate / fastNonGate say, as the
      // reference sets fastReaction only for tolerance 1e-4 with mod 1
      // (elec/TT06Dev_Reaction.cc:110-114)
      const bool exact = p->tolerance <= 0.0;
      if (!exact) check_configuration_b(p->tolerance, p->mod, p->fastGate, p->fastNonGate);
      CDH_REQUIRE(p->dt > 0, "cdh_tt06_create: dt must be positive");
      auto r = std::make_unique<cdh_tt06>();
      r->n = numPoints;
      r->stride = (numPoints + 63) / 64 * 64;
      r->exact = exact;
      tt06_constants(&r->K);
      std::memcpy(&r->ct, &p->celltype, sizeof(CellType));
      std::vector<double> coef;
      if (exact) {
         TT06ExactConst E;
         tt06_exact_constants(&E, p->mod);
         r->Ed.alloc(1);
         r->Ed.upload(&E, 1);
         coef.assign(kFitCoefTotal, 0.0);
      } else {
         coef = scaled_fit_coefficients(p->fit_l, p->fit_m, p->fit_coef, p->dt);
      }
      r->coef.alloc(coef.size());
      r->coef.upload(coef.data(), coef.size());
      r->Kd.alloc(1);
      r->Kd.upload(&r->K, 1);
      ParamSet ps = make_param_set(r->ct, p->s_kr, p->s_cal, p->s_na);
      r->params.alloc(1);
      r->params.upload(&ps, 1);
      // configuration B reproduces the v2 loops' per-group bounds and f-gate
      // rounding per cell; the generic gate loop of configuration A visits
      // every cell alike
      std::vector<uint8_t> fl(r->stride, kFlagGatesFixed), fixed(r->stride, kFlagGatesFixed);
      for (int i = 0; i < numPoints; i++) fl[i] = exact ? 0 : gate_flags(i, numPoints);
      r->flags.alloc(r->stride);
      r->flags.upload(fl.data(), fl.size());
      r->flagsAllFixed.alloc(r->stride);
      r->flagsAllFixed.upload(fixed.data(), fixed.size());
      // states in 64-cell tiles (state s of cell i at blk_index(s, i, 19)): a
      // wavefront's 19 rows are one contiguous 9.5 KB run
      const std::vector<double> st = rest_state_soa(r->ct, numPoints, r->stride, r->K.c9);
      std::vector<double> tiled(st.size());
      for (int s = 0; s < N_STATE; s++)
         for (int64_t i = 0; i < r->stride; i++) tiled[blk_index(s, i, N_STATE)] = st[s * r->stride + i];
      r->state.alloc(tiled.size());
      r->state.upload(tiled.data(), tiled.size());
      if (!exact) {
         // the pipelined kernel's grid: as many workgroups as are resident at once
         int dev = 0, ncu = 0, per_cu = 0;
         CDH_HIP(hipGetDevice(&dev));
         CDH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
         CDH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tt06_calc_pipe_f64,
                                                              kBlock, 0));
         CDH_REQUIRE(ncu > 0 && per_cu > 0, "cdh_tt06_create: no resident workgroups");
         r->pipe_blocks = (int64_t)ncu * per_cu;
         r->sink.alloc((N_STATE + 1) * kBlk);
      }
      *out = r.release();
   });
}

int cdh_tt06_destroy(cdh_tt06 *r)
{
   delete r;
   return CDH_OK;
}

int cdh_tt06_calc(cdh_tt06 *r, double dt, const int *indexArray, int n, const double *Vm,
                  const double *iStim, double *dVm, void *stream)
{
   (void)iStim;
   return guard([&] {
      CDH_REQUIRE(r && Vm && dVm, "cdh_tt06_calc: null argument");
      CDH_REQUIRE(n == r->n, "cdh_tt06_calc: n differs from numPoints");
      if (n == 0) return;
      launch_calc(r, reaction_args(r, dt, indexArray, n, Vm, dVm, r->flags.p), S(stream));
   });
}

int cdh_tt06_update_nongate(cdh_tt06 *r, double dt, const int *indexArray, int n,
                            const double *Vm, double *dVR, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(r && Vm && dVR && n == r->n, "cdh_tt06_update_nongate: bad arguments");
      if (n == 0) return;
      // the calc kernel with every gate held fixed is the non-gate update
      launch_calc(r, reaction_args(r, dt, indexArray, n, Vm, dVR, r->flagsAllFixed.p),
                  S(stream));
   });
}

int cdh_tt06_update_gate(cdh_tt06 *r, double dt, const int *indexArray, int n, const double *Vm,
                         void *stream)
{
   return guard([&] {
      CDH_REQUIRE(r && Vm && n == r->n, "cdh_tt06_update_gate: bad arguments");
      if (n == 0) return;
      ReactionArgs a = reaction_args(r, dt, indexArray, n, Vm, nullptr, r->flags.p);
      hipLaunchKernelGGL(r->exact ? k_tt06_gate_exact_f64 : k_tt06_gate_f64,
                         dim3(grid_for(n, kBlock)), dim3(kBlock), 0, S(stream), a);
      CDH_LAUNCH_CHECK();
   });
}

int cdh_tt06_initialize_membrane_voltage(cdh_tt06 *r, const int *indexArray, int n, double *Vm,
                                         void *stream)
{
   return guard([&] {
      CDH_REQUIRE(r && Vm && n == r->n, "cdh_tt06_initialize_membrane_voltage: bad arguments");
      if (n == 0) return;
      if (indexArray) {
         hipLaunchKernelGGL(k_set_list_f64, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, S(stream),
                            n, indexArray, r->ct.Vm0, Vm);
      } else {
         hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, S(stream), n,
                            r->ct.Vm0, Vm);
      }
      CDH_LAUNCH_CHECK();
   });
}

int cdh_tt06_checkpoint_field(int i, const char **name, const char **unit, int *handle)
{
   if (i < 0 || i >= N_STATE) {
      set_last_error("cdh_tt06_checkpoint_field: index out of range");
      return CDH_ERR_ARG;
   }
   if (name) *name = kFieldName[i];
   if (unit) *unit = kFieldUnit[i];
   if (handle) *handle = i;
   return CDH_OK;
}

int cdh_tt06_var_handle(const char *name)
{
   if (!name) return -1;
   for (int i = 0; i < N_STATE; i++)
      if (std::strcmp(kFieldName[i], name) == 0) return i;
   return -1;
}

int cdh_tt06_get_value(cdh_tt06 *r, int iCell, int handle, double *value)
{
   return guard([&] {
      CDH_REQUIRE(r && value && iCell >= 0 && iCell < r->n && handle >= 0 && handle < N_STATE,
                  "cdh_tt06_get_value: bad arguments");
      CDH_HIP(hipDeviceSynchronize());
      r->state.download(value, 1, blk_index(handle, iCell, N_STATE));
   });
}

int cdh_tt06_set_value(cdh_tt06 *r, int iCell, int handle, double value)
{
   return guard([&] {
      CDH_REQUIRE(r && iCell >= 0 && iCell < r->n && handle >= 0 && handle < N_STATE,
                  "cdh_tt06_set_value: bad arguments");
      CDH_HIP(hipDeviceSynchronize());
      r->state.upload(&value, 1, blk_index(handle, iCell, N_STATE));
   });
}

int cdh_tt06_get_state(cdh_tt06 *r, double *host)
{
   return guard([&] {
      CDH_REQUIRE(r && host, "cdh_tt06_get_state: null argument");
      CDH_HIP(hipDeviceSynchronize());
      std::vector<double> tiled((size_t)N_STATE * r->stride);
      r->state.download(tiled.data(), tiled.size());
      for (int s = 0; s < N_STATE; s++)
         for (int64_t i = 0; i < r->n; i++) host[(size_t)s * r->n + i] = tiled[blk_index(s, i, N_STATE)];
   });
}

int cdh_tt06_set_state(cdh_tt06 *r, const double *host)
{
   return guard([&] {
      CDH_REQUIRE(r && host, "cdh_tt06_set_state: null argument");
      CDH_HIP(hipDeviceSynchronize());
      std::vector<double> tiled((size_t)N_STATE * r->stride);
      r->state.download(tiled.data(), tiled.size());
      for (int s = 0; s < N_STATE; s++)
         for (int64_t i = 0; i < r->n; i++) tiled[blk_index(s, i, N_STATE)] = host[(size_t)s * r->n + i];
      r->state.upload(tiled.data(), tiled.size());
   });
}

// ---------------------------------------------------------------------------
// Diffusion
// ---------------------------------------------------------------------------
int cdh_fgr_create(const cdh_fgr_params *p, cdh_fgr **out)
{
   return guard([&] {
      CDH_REQUIRE(p && out && p->gid && p->sigma && p->cellType, "cdh_fgr_create: null argument");
      CDH_REQUIRE(p->nLocal >= 0 && p->nRemote >= 0, "cdh_fgr_create: negative sizes");
      auto d = std::make_unique<cdh_fgr>();
      const int64_t ntot = p->nLocal + p->nRemote;
      d->nLocal = p->nLocal;
      d->nRemote = p->nRemote;
      d->stride = (p->nLocal + 63) / 64 * 64;
      d->scale = p->diffusionScale;
      d->geom = bounding_box(p->nx, p->ny, p->nz, p->gid, p->nLocal, p->dx, p->dy, p->dz);
      const BoxGeom &g = d->geom;
      const size_t nbox = (size_t)g.bx * g.by * g.bz;
      CDH_REQUIRE(nbox < (size_t)UINT32_MAX, "cdh_fgr_create: box too large for 32-bit indices");
      // The box is private to this object, so it is laid out x fastest, as
      // the callers' gid order runs: a wavefront's 64 consecutive cells then
      // gather from consecutive box points (the reference's Array3d is z
      // fastest; only the offsets differ, not the weights or the sums).
      auto box_index_x = [&](int64_t gid) -> int64_t {
         const int64_t x = gid % p->nx, y = (gid / p->nx) % p->ny;
         const int64_t z = gid / ((int64_t)p->nx * p->ny);
         const int64_t lx = x - g.x0, ly = y - g.y0, lz = z - g.z0;
         if (lx < 0 || ly < 0 || lz < 0 || lx >= g.bx || ly >= g.by || lz >= g.bz) return -1;
         return lx + (int64_t)g.bx * (ly + (int64_t)g.by * lz);
      };
      int64_t off[19];
      fgr_offsets_strided(1, g.bx, (int64_t)g.bx * g.by, off);
      // the stencil takes each row's x-1 and x+1 values from the
      // neighbouring lanes (row_shift.hpp): they sit one box index either
      // side of the row centre's
      for (int r = 0; r < 5; r++)
         CDH_REQUIRE(off[kRowsX.back[r]] == off[kRowsX.centre[r]] - 1 &&
                         off[kRowsX.on[r]] == off[kRowsX.centre[r]] + 1,
                     "cdh_fgr_create: stencil rows are not along x");
      std::vector<int64_t> bidx(ntot);
      std::vector<uint8_t> tissue(nbox, 0);
      std::vector<double> sig(6 * nbox, 0.0);
      for (int64_t c = 0; c < ntot; c++) {
         int64_t ib = box_index_x(p->gid[c]);
         CDH_REQUIRE(ib >= 0, "cdh_fgr_create: a remote cell lies outside the padded box");
         bidx[c] = ib;
         tissue[ib] = p->cellType[c] > 0 ? 1 : 0;
         std::memcpy(&sig[6 * ib], p->sigma + 6 * c, 6 * sizeof(double));
      }
      std::vector<uint32_t> b32(ntot);
      for (int64_t c = 0; c < ntot; c++) b32[c] = (uint32_t)bidx[c];
      d->bi.alloc(ntot ? ntot : 1);
      d->bi.upload(b32.data(), ntot);
      d->box.alloc(nbox);
      d->box.zero();
      // precomputeCoefficients on the device; padding lanes keep zero weights
      d->A0.alloc(d->stride ? d->stride : 1);
      d->A0.zero();
      d->Aw.alloc(18 * (d->stride ? d->stride : 1));
      d->Aw.zero();
      if (p->nLocal > 0) {
         DevBuf<uint8_t> dT(nbox);
         DevBuf<double> dS(6 * nbox);
         DevBuf<int64_t> dB(p->nLocal);
         dT.upload(tissue.data(), nbox);
         dS.upload(sig.data(), 6 * nbox);
         dB.upload(bidx.data(), p->nLocal);
         // weights in 64-cell tiles, in the stencil's load groups
         // (grp_index18): a wavefront's 18 weights are one contiguous 4.5 KB
         // run that five load instructions bring in
         device_fgr_weights(nullptr, dT.p, dS.p, p->nLocal, dB.p, fgr::weight_geom(g, off),
                            d->A0.p, d->Aw.p, d->stride, kWeightsGrouped);
         CDH_HIP(hipDeviceSynchronize());
      }
      StencilArgs &a = d->args;
      a.n = (int)p->nLocal;
      a.bi = d->bi.p;
      a.box = d->box.p;
      a.Aw = d->Aw.p;
      for (int k = 0; k < 19; k++) a.off[k] = (int)off[k];
      a.scale = d->scale;
      a.dVm = nullptr;
      a.groups_per_xcd = (int)((grid_for(p->nLocal, kBlock) + 7) / 8);
      *out = d.release();
   });
}

int cdh_fgr_destroy(cdh_fgr *d)
{
   delete d;
   return CDH_OK;
}

int cdh_fgr_update_local_voltage(cdh_fgr *d, const double *VmLocal, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(d && VmLocal, "cdh_fgr_update_local_voltage: null argument");
      if (d->nLocal == 0) return;
      scatter(d->nLocal, d->bi.p, VmLocal, d->box.p, S(stream));
   });
}

int cdh_fgr_update_remote_voltage(cdh_fgr *d, const double *VmRemote, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(d, "cdh_fgr_update_remote_voltage: null argument");
      if (d->nRemote == 0) return;
      CDH_REQUIRE(VmRemote, "cdh_fgr_update_remote_voltage: null argument");
      scatter(d->nRemote, d->bi.p + d->nLocal, VmRemote, d->box.p, S(stream));
   });
}

int cdh_fgr_calc(cdh_fgr *d, double *dVm, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(d && dVm, "cdh_fgr_calc: null argument");
      if (d->nLocal == 0) return;
      StencilArgs a = d->args;
      a.dVm = dVm;
      hipLaunchKernelGGL(k_fgr_calc_f64, dim3(8 * a.groups_per_xcd), dim3(kBlock), 0,
                         S(stream), a);
      CDH_LAUNCH_CHECK();
   });
}

double cdh_fgr_diffusion_scale(const cdh_fgr *d) { return d ? d->scale : 0.0; }

// ---------------------------------------------------------------------------
// Loop pieces
// ---------------------------------------------------------------------------
int cdh_fill(int64_t n, double value, double *x, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(n >= 0 && n < INT32_MAX && (x || n == 0), "cdh_fill: bad arguments");
      if (n == 0) return;
      hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, S(stream), (int)n,
                         value, x);
      CDH_LAUNCH_CHECK();
   });
}

int cdh_set_list(int64_t n, const int *list, double value, double *x, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(n >= 0 && n < INT32_MAX && ((list && x) || n == 0), "cdh_set_list: bad arguments");
      if (n == 0) return;
      hipLaunchKernelGGL(k_set_list_f64, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, S(stream),
                         (int)n, list, value, x);
      CDH_LAUNCH_CHECK();
   });
}

int cdh_integrate(int64_t n, double dt, double *Vm, const double *dVmR, const double *dVmD,
                  const double *iStim, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(n >= 0 && n < INT32_MAX && ((Vm && dVmR && dVmD && iStim) || n == 0),
                  "cdh_integrate: bad arguments");
      if (n == 0) return;
      hipLaunchKernelGGL(k_integrate_f64, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, S(stream),
                         (int)n, dt, Vm, dVmR, dVmD, iStim);
      CDH_LAUNCH_CHECK();
   });
}

int cdh_activation_sensor_create(int64_t n, double threshold, cdh_sensor **out)
{
   return guard([&] {
      CDH_REQUIRE(out && n >= 0 && n < INT32_MAX, "cdh_activation_sensor_create: bad arguments");
      auto s = std::make_unique<cdh_sensor>();
      s->n = n;
      s->threshold = threshold;
      const size_t m = n ? n : 1;
      std::vector<double> neg(m, -1.0);
      s->tact0.alloc(m);
      s->tact0.upload(neg.data(), m);
      s->tact.alloc(m);
      s->tact.upload(neg.data(), m);
      s->trec.alloc(m);
      s->trec.upload(neg.data(), m);
      s->active.alloc(m);
      s->active.zero();
      s->activated.alloc(m);
      s->activated.zero();
      s->counts.alloc(4);
      s->counts.zero();
      *out = s.release();
   });
}

int cdh_activation_sensor_eval(cdh_sensor *s, double time, const double *Vm, void *stream)
{
   return guard([&] {
      CDH_REQUIRE(s && (Vm || s->n == 0), "cdh_activation_sensor_eval: bad arguments");
      if (s->n == 0) return;
      SensorArgs a;
      a.n = (int)s->n;
      a.enabled = 1;
      a.first_eval = s->evaluated == 0;
      a.time = time;
      a.threshold = s->threshold;
      a.tact0 = s->tact0.p;
      a.tact = s->tact.p;
      a.trec = s->trec.p;
      a.active = s->active.p;
      a.activated = s->activated.p;
      a.apd = s->apd;
      a.apd_phase = s->apd_phase.p;
      a.vmax = s->vmax.p;
      a.t90 = s->t90.p;
      a.counts = s->counts.p;
      if (aligned16(Vm))
         hipLaunchKernelGGL(k_sensor4_f64, dim3(grid_for((s->n + 3) / 4, kBlock)), dim3(kBlock), 0,
                            S(stream), a, Vm, (const double *)s->vrest.p);
      else
         hipLaunchKernelGGL(k_sensor_f64, dim3(grid_for(s->n, kBlock)), dim3(kBlock), 0,
                            S(stream), a, Vm, (const double *)s->vrest.p);
      CDH_LAUNCH_CHECK();
      s->evaluated = 1;
   });
}

int cdh_activation_sensor_counts(cdh_sensor *s, int64_t *n_tact0, int64_t *n_tact,
                                  int64_t *n_trec)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_activation_sensor_counts: null argument");
      uint32_t c[4] = {0, 0, 0, 0};
      CDH_HIP(hipDeviceSynchronize());
      if (s->n > 0) s->counts.download(c, 4);
      if (n_tact0) *n_tact0 = c[0];
      if (n_tact) *n_tact = c[1];
      if (n_trec) *n_trec = c[2];
   });
}

int cdh_activation_sensor_get(cdh_sensor *s, double *tact0, double *tact, double *trec)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_activation_sensor_get: null argument");
      CDH_HIP(hipDeviceSynchronize());
      if (tact0) s->tact0.download(tact0, s->n);
      if (tact) s->tact.download(tact, s->n);
      if (trec) s->trec.download(trec, s->n);
   });
}

int cdh_activation_sensor_enable_apd(cdh_sensor *s, const double *vrest)
{
   return guard([&] {
      CDH_REQUIRE(s && (vrest || s->n == 0), "cdh_activation_sensor_enable_apd: null argument");
      CDH_REQUIRE(!s->evaluated, "cdh_activation_sensor_enable_apd: sensor already evaluated");
      const size_t m = s->n ? s->n : 1;
      s->apd_phase.alloc(m);
      s->apd_phase.zero();
      s->vmax.alloc(m);
      s->vmax.zero();
      std::vector<double> neg(m, -1.0);
      s->t90.alloc(m);
      s->t90.upload(neg.data(), m);
      s->vrest.alloc(m);
      s->vrest.upload(vrest, s->n);
      s->apd = 1;
   });
}

int cdh_activation_sensor_get_t90(cdh_sensor *s, double *t90)
{
   return guard([&] {
      CDH_REQUIRE(s && t90, "cdh_activation_sensor_get_t90: null argument");
      CDH_REQUIRE(s->apd, "cdh_activation_sensor_get_t90: APD90 not enabled");
      CDH_HIP(hipDeviceSynchronize());
      s->t90.download(t90, s->n);
   });
}

int cdh_activation_sensor_destroy(cdh_sensor *s)
{
   delete s;
   return CDH_OK;
}

}  // extern "C"
