// sim.hip -- the monodomain engine: simulationLoop (elec/simulationLoop.cc:
// 150-243) with the reaction, diffusion, stimuli and activation sensors
// device resident.
//
// Cells are held in box order (x-major, z fastest) so that a cell's stencil
// neighbours are near it in memory; the input order is kept in `perm` for
// the host accessors.  The reaction's per-cell gate flags are derived from
// each cell's index in its reaction group in the input order, as the
// reference's ReactionManager groups cells (elec/ReactionManager.cc:218-261).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <tuple>
#include <vector>

#include "../../include/cardioid_hip.h"
#include "fgr_weights.hpp"
#include "hip_util.hpp"
#include "kernels.hpp"
#include "setup.hpp"
#include "tt06_f64.hpp"

namespace cdh {

__global__ void k_integrate_sensor_f64(int n, double dt, double *Vm, const double *dVmR,
                                       const double *dVmD, const double *iStim, SensorArgs s,
                                       const ParamSet *params, const uint8_t *pid);

}  // namespace cdh

using namespace cdh;

struct cdh_sim {
   cdh_sim_params p{};
   int64_t n = 0, stride = 0;
   BoxGeom geom{};
   int mode = CDH_MODE_FP64_FUSED;
   int variant = 3;                      // fused kernel occupancy target (waves/SIMD)
   hipStream_t stream = nullptr;
   hipEvent_t ev0 = nullptr, ev1 = nullptr;
   TT06Const K{};
   std::vector<int64_t> perm;            // engine index -> input index
   std::vector<cdh_box_stim> stim;
   std::vector<std::vector<int>> stimList;   // engine indices per stimulus
   DevBuf<double> coef;
   DevBuf<ParamSet> params;
   DevBuf<uint8_t> pid, flags, stimMask;
   // per-cell multi-field arrays in 64-cell blocks (blk_index), stride cells
   DevBuf<double> state;              // FP64 modes: 19 fields
   DevBuf<float> st32;                // mixed mode: 17 fields
   DevBuf<double> st64;               // mixed mode: 2 fields (Ca_i, Na_i)
   std::vector<double> K0cell;        // mixed mode: rest K_i per engine cell (host)
   DevBuf<uint32_t> bi;
   DevBuf<double> A0;
   DevBuf<float> Aw;
   DevBuf<double> box[2];
   int cur = 0;
   DevBuf<int> stimListDev;
   std::vector<int64_t> stimListOff;
   // unfused mode
   DevBuf<double> Vm, dVmR, dVmD, iStim;
   // sensors
   DevBuf<double> tact0, tact, trec;
   DevBuf<uint8_t> apdPhase;
   DevBuf<double> vmax, t90;
   int off[19] = {0};
   double time = 0.0;
   int loop = 0;
   int64_t bytes = 0;

   ~cdh_sim()
   {
      if (ev0) (void)hipEventDestroy(ev0);
      if (ev1) (void)hipEventDestroy(ev1);
      if (stream) (void)hipStreamDestroy(stream);
   }
};

namespace {

constexpr int kBlock = 256;

// mixed-mode FP32 slots (tt06_mixed.hpp order) and the reference slot each holds
constexpr int mixed_n32 = 17;
int mixed_slot_source(int q)
{
   static const int src[mixed_n32] = {S_CA_SS, S_CA_SR, S_R_PRIME, S_FCASS, -1 /* K_i - K0 */,
                                      S_M, S_H, S_J, S_XR1, S_XR2, S_XS, S_R, S_D, S_F, S_F2,
                                      S_JL, S_S};
   return src[q];
}

// Vm per engine cell from the current box
std::vector<double> engine_vm(cdh_sim *s);

// Host copies of the engine's multi-field arrays are [nq][S] (S = stride, a
// multiple of 64); on the device they are in 64-cell blocks (blk_index).
template <class T>
void upload_blocked(DevBuf<T> &d, const std::vector<T> &soa, int nq, int64_t S)
{
   std::vector<T> b(soa.size());
   for (int q = 0; q < nq; q++)
      for (int64_t k = 0; k < S; k++) b[blk_index(q, k, nq)] = soa[q * S + k];
   d.upload(b.data(), b.size());
}

template <class T>
void download_blocked(DevBuf<T> &d, std::vector<T> &soa, int nq, int64_t S)
{
   std::vector<T> b(soa.size());
   d.download(b.data(), b.size());
   for (int q = 0; q < nq; q++)
      for (int64_t k = 0; k < S; k++) soa[q * S + k] = b[blk_index(q, k, nq)];
}

// all 19 states per engine cell, [19][stride], the reference's layout
// (slot 1 = dVK_i = K_i/c9 + Vm)
std::vector<double> state19(cdh_sim *s)
{
   const int64_t n = s->n, S = s->stride;
   std::vector<double> st((size_t)N_STATE * S, 0.0);
   if (s->mode != CDH_MODE_MIXED_FUSED) {
      download_blocked(s->state, st, N_STATE, S);
      return st;
   }
   std::vector<float> s32((size_t)mixed_n32 * S);
   std::vector<double> s64((size_t)2 * S);
   download_blocked(s->st32, s32, mixed_n32, S);
   download_blocked(s->st64, s64, 2, S);
   std::vector<double> vm = engine_vm(s);
   for (int64_t k = 0; k < n; k++) {
      for (int q = 0; q < mixed_n32; q++) {
         const int d = mixed_slot_source(q);
         if (d >= 0) st[d * S + k] = s32[q * S + k];
      }
      st[S_CA_I * S + k] = s64[k];
      st[S_NA_I * S + k] = s64[S + k];
      const double K_i = s->K0cell[k] + (double)s32[4 * S + k];
      st[S_DVK_I * S + k] = K_i / s->K.c9 + vm[k];
   }
   return st;
}

void set_state19(cdh_sim *s, const std::vector<double> &st, const std::vector<double> &vm)
{
   const int64_t n = s->n, S = s->stride;
   if (s->mode != CDH_MODE_MIXED_FUSED) {
      upload_blocked(s->state, st, N_STATE, S);
      return;
   }
   std::vector<float> s32((size_t)mixed_n32 * S, 0.0f);
   std::vector<double> s64((size_t)2 * S, 0.0);
   for (int64_t k = 0; k < n; k++) {
      for (int q = 0; q < mixed_n32; q++) {
         const int d = mixed_slot_source(q);
         if (d >= 0) s32[q * S + k] = (float)st[d * S + k];
      }
      const double K_i = s->K.c9 * (st[S_DVK_I * S + k] - vm[k]);
      s32[4 * S + k] = (float)(K_i - s->K0cell[k]);
      s64[k] = st[S_CA_I * S + k];
      s64[S + k] = st[S_NA_I * S + k];
   }
   upload_blocked(s->st32, s32, mixed_n32, S);
   upload_blocked(s->st64, s64, 2, S);
}

// PeriodicPulse::eval (elec/PeriodicPulse.cc:22-28)
double pulse(const cdh_box_stim &b, double time)
{
   double t = time - std::floor(time / b.period) * b.period;
   if (t >= b.phase && t < b.phase + b.duration) return b.amplitude;
   return 0.0;
}

// stimulus values of the step starting at `time` (Stimulus::stim, Stimulus.hh:21-27)
StimStep stim_step(const cdh_sim *s, double time)
{
   StimStep st{};
   st.nstim = (int)s->stim.size();
   st.on = 0;
   for (int k = 0; k < st.nstim; k++) {
      const cdh_box_stim &b = s->stim[k];
      st.value[k] = 0.0;
      if (!(time > b.t0 && time < b.tf)) continue;
      const double v = pulse(b, time);
      if (v == 0) continue;
      st.value[k] = v;
      st.on |= 1u << k;
   }
   return st;
}

template <class T>
void track(cdh_sim *s, DevBuf<T> &b, size_t count)
{
   b.alloc(count);
   s->bytes += (int64_t)(count * sizeof(T));
}

void step_unfused(cdh_sim *s);

SensorArgs sensor_args(cdh_sim *s)
{
   SensorArgs a{};
   a.n = (int)s->n;
   a.enabled = s->p.sensors;
   a.first_eval = s->loop == 1;
   a.time = s->time;
   a.threshold = s->p.threshold;
   a.tact0 = s->tact0.p;
   a.tact = s->tact.p;
   a.trec = s->trec.p;
   a.active = nullptr;
   a.activated = nullptr;
   a.apd = (s->p.sensors & 2) ? 1 : 0;
   a.apd_phase = s->apdPhase.p;
   a.vmax = s->vmax.p;
   a.t90 = s->t90.p;
   return a;
}

void step_fused(cdh_sim *s)
{
   FusedArgs a;
   a.K = s->K;
   a.coef = s->coef.p;
   a.params = s->params.p;
   a.pid = s->pid.p;
   a.flags = s->flags.p;
   a.state = s->state.p;
   a.stride = s->stride;
   a.n = (int)s->n;
   const int nblock = grid_for(s->n, kBlock);
   a.nblock_per_xcd = (nblock + 7) / 8;
   a.bi = s->bi.p;
   a.box_cur = s->box[s->cur].p;
   a.box_next = s->box[s->cur ^ 1].p;
   a.A0 = s->A0.p;
   a.Aw = s->Aw.p;
   for (int k = 0; k < 19; k++) a.off[k] = s->off[k];
   a.scale = s->p.diffusionScale;
   a.dt = s->p.dt;
   a.gate_mask = 0xFFFu;
   a.stim = stim_step(s, s->time);
   a.stim_any = a.stim.on != 0;
   a.stim_mask = s->stimMask.p;
   // sensors see the time and loop after the step
   s->time += s->p.dt;
   ++s->loop;
   a.sens = sensor_args(s);
   if (a.n > 0) {
      const dim3 grid(8 * a.nblock_per_xcd), block(kBlock);
      switch (s->variant) {
      case 2: hipLaunchKernelGGL(k_step_fused_f64<2>, grid, block, 0, s->stream, a); break;
      case 4: hipLaunchKernelGGL(k_step_fused_f64<4>, grid, block, 0, s->stream, a); break;
      default: hipLaunchKernelGGL(k_step_fused_f64<3>, grid, block, 0, s->stream, a); break;
      }
      CDH_LAUNCH_CHECK();
   }
   s->cur ^= 1;
}

void step_mixed(cdh_sim *s)
{
   MixedArgs a;
   a.K = s->K;
   a.coef = s->coef.p;
   a.params = s->params.p;
   a.pid = s->pid.p;
   a.flags = s->flags.p;
   a.st32 = s->st32.p;
   a.st64 = s->st64.p;
   a.stride = s->stride;
   a.n = (int)s->n;
   const int nblock = grid_for(s->n, kBlock);
   a.nblock_per_xcd = (nblock + 7) / 8;
   a.bi = s->bi.p;
   a.box_cur = s->box[s->cur].p;
   a.box_next = s->box[s->cur ^ 1].p;
   a.Aw = s->Aw.p;
   for (int k = 0; k < 19; k++) a.off[k] = s->off[k];
   a.scale = s->p.diffusionScale;
   a.dt = s->p.dt;
   a.gate_mask = 0xFFFu;
   a.stim = stim_step(s, s->time);
   a.stim_any = a.stim.on != 0;
   a.stim_mask = s->stimMask.p;
   s->time += s->p.dt;
   ++s->loop;
   a.sens = sensor_args(s);
   if (a.n > 0) {
      const dim3 grid(8 * a.nblock_per_xcd), block(kBlock);
      switch (s->variant) {
      case 4: hipLaunchKernelGGL(k_step_fused_mixed<4>, grid, block, 0, s->stream, a); break;
      case 5: hipLaunchKernelGGL(k_step_fused_mixed<5>, grid, block, 0, s->stream, a); break;
      case 6: hipLaunchKernelGGL(k_step_fused_mixed<6>, grid, block, 0, s->stream, a); break;
      default: hipLaunchKernelGGL(k_step_fused_mixed<3>, grid, block, 0, s->stream, a); break;
      }
      CDH_LAUNCH_CHECK();
   }
   s->cur ^= 1;
}

void step(cdh_sim *s)
{
   if (s->mode == CDH_MODE_FP64_FUSED) step_fused(s);
   else if (s->mode == CDH_MODE_MIXED_FUSED) step_mixed(s);
   else step_unfused(s);
}

void step_unfused(cdh_sim *s)
{
   const int n = (int)s->n;
   const dim3 g(grid_for(n, kBlock)), b(kBlock);
   if (n == 0) {
      s->time += s->p.dt;
      ++s->loop;
      return;
   }
   // iStim = 0; each active stimulus assigns its value on its cells
   hipLaunchKernelGGL(k_fill_f64, g, b, 0, s->stream, n, 0.0, s->iStim.p);
   StimStep st = stim_step(s, s->time);
   for (int k = 0; k < st.nstim; k++) {
      if (!((st.on >> k) & 1u)) continue;
      const int cnt = (int)(s->stimListOff[k + 1] - s->stimListOff[k]);
      if (cnt == 0) continue;
      hipLaunchKernelGGL(k_set_list_f64, dim3(grid_for(cnt, kBlock)), b, 0, s->stream, cnt,
                         s->stimListDev.p + s->stimListOff[k], st.value[k], s->iStim.p);
   }
   // reaction
   ReactionArgs r;
   r.K = s->K;
   r.coef = s->coef.p;
   r.params = s->params.p;
   r.pid = s->pid.p;
   r.flags = s->flags.p;
   r.state = s->state.p;
   r.stride = s->stride;
   r.n = n;
   r.index = nullptr;
   r.Vm = s->Vm.p;
   r.dVm = s->dVmR.p;
   r.dt = s->p.dt;
   r.gate_mask = 0xFFFu;
   hipLaunchKernelGGL(k_tt06_calc_f64<true>, g, b, 0, s->stream, r);
   // diffusion: scatter Vm into the box, then the stencil
   hipLaunchKernelGGL(k_scatter_f64, g, b, 0, s->stream, n, s->bi.p, s->Vm.p, s->box[0].p);
   StencilArgs d;
   d.n = n;
   d.stride = s->stride;
   d.bi = s->bi.p;
   d.box = s->box[0].p;
   d.A0 = s->A0.p;
   d.Aw = s->Aw.p;
   for (int k = 0; k < 19; k++) d.off[k] = s->off[k];
   d.scale = s->p.diffusionScale;
   d.dVm = s->dVmD.p;
   hipLaunchKernelGGL(k_fgr_calc_f64<true>, g, b, 0, s->stream, d);
   // integrator and sensors
   s->time += s->p.dt;
   ++s->loop;
   hipLaunchKernelGGL(k_integrate_sensor_f64, g, b, 0, s->stream, n, s->p.dt, s->Vm.p, s->dVmR.p,
                      s->dVmD.p, s->iStim.p, sensor_args(s), s->params.p, s->pid.p);
   CDH_LAUNCH_CHECK();
}

std::vector<double> engine_vm(cdh_sim *s)
{
   const int64_t n = s->n;
   std::vector<double> v(n);
   if (n == 0) return v;
   if (s->mode != CDH_MODE_FP64_UNFUSED) {
      const BoxGeom &g = s->geom;
      std::vector<double> box((size_t)g.bx * g.by * g.bz);
      s->box[s->cur].download(box.data(), box.size());
      std::vector<uint32_t> bi(n);
      s->bi.download(bi.data(), n);
      for (int64_t k = 0; k < n; k++) v[k] = box[bi[k]];
   } else {
      s->Vm.download(v.data(), n);
   }
   return v;
}

}  // namespace

extern "C" {

int cdh_sim_create(const cdh_sim_params *p, int nstim, const cdh_box_stim *stim, cdh_sim **out)
{
   return guard([&] {
      CDH_REQUIRE(p && out, "cdh_sim_create: null argument");
      CDH_REQUIRE(p->ncell >= 0 && p->ncell < INT32_MAX, "cdh_sim_create: ncell out of range");
      CDH_REQUIRE(p->ncell == 0 || (p->gid && p->celltype && p->sigma), "cdh_sim_create: null arrays");
      CDH_REQUIRE(p->ncelltype >= 1 && p->ncelltype <= kMaxCellTypes && p->celltypes,
                  "cdh_sim_create: 1..8 cell types required");
      CDH_REQUIRE(nstim >= 0 && nstim <= kMaxStim && (nstim == 0 || stim),
                  "cdh_sim_create: at most 8 stimuli");
      CDH_REQUIRE(p->mode == CDH_MODE_FP64_UNFUSED || p->mode == CDH_MODE_FP64_FUSED ||
                      p->mode == CDH_MODE_MIXED_FUSED,
                  "cdh_sim_create: unknown mode");
      CDH_REQUIRE(p->dt > 0, "cdh_sim_create: dt must be positive");
      auto s = std::make_unique<cdh_sim>();
      s->p = *p;
      s->mode = p->mode;
      if (p->mode == CDH_MODE_MIXED_FUSED) {
         CDH_REQUIRE(p->variant == 0 || (p->variant >= 3 && p->variant <= 6),
                     "cdh_sim_create: mixed variant must be 0 (default) or 3..6");
         s->variant = p->variant ? p->variant : 3;
      } else {
         CDH_REQUIRE(p->variant == 0 || (p->variant >= 2 && p->variant <= 4),
                     "cdh_sim_create: variant must be 0 (default) or 2..4");
         s->variant = p->variant ? p->variant : 3;
      }
      const int64_t n = p->ncell;
      s->n = n;
      s->stride = std::max<int64_t>(64, (n + 63) / 64 * 64);
      tt06_constants(&s->K);
      CDH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      CDH_HIP(hipEventCreate(&s->ev0));
      CDH_HIP(hipEventCreate(&s->ev1));

      // box and engine order
      s->geom = bounding_box(p->nx, p->ny, p->nz, p->gid, n, p->dx, p->dy, p->dz);
      const BoxGeom &g = s->geom;
      const size_t nbox = n ? (size_t)g.bx * g.by * g.bz : 1;
      CDH_REQUIRE(nbox < (size_t)UINT32_MAX, "cdh_sim_create: box too large for 32-bit indices");
      std::vector<int64_t> bidxIn(n);
      for (int64_t c = 0; c < n; c++) {
         CDH_REQUIRE(p->celltype[c] < p->ncelltype, "cdh_sim_create: cell type index out of range");
         bidxIn[c] = box_index(g, p->nx, p->ny, p->gid[c]);
      }
      s->perm.resize(n);
      std::iota(s->perm.begin(), s->perm.end(), 0);
      std::sort(s->perm.begin(), s->perm.end(),
                [&](int64_t a, int64_t b) { return bidxIn[a] < bidxIn[b]; });
      for (int64_t k = 1; k < n; k++)
         CDH_REQUIRE(bidxIn[s->perm[k]] != bidxIn[s->perm[k - 1]], "cdh_sim_create: duplicate gid");

      // gate flags by reaction group (input order), parameter sets
      std::vector<int64_t> groupCount(p->ncelltype, 0), groupPos(n);
      for (int64_t c = 0; c < n; c++) groupPos[c] = groupCount[p->celltype[c]]++;
      std::vector<uint8_t> flIn(n);
      for (int64_t c = 0; c < n; c++)
         flIn[c] = p->gate_tail_rule ? gate_flags(groupPos[c], groupCount[p->celltype[c]]) : 0;
      std::map<std::tuple<int, double, double, double>, int> psetId;
      std::vector<ParamSet> psets;
      std::vector<uint8_t> pidIn(n);
      std::vector<CellType> cts(p->ncelltype);
      std::memcpy(cts.data(), p->celltypes, sizeof(CellType) * p->ncelltype);
      for (int64_t c = 0; c < n; c++) {
         const double skr = p->scales ? p->scales[3 * c] : 1.0;
         const double scal = p->scales ? p->scales[3 * c + 1] : 1.0;
         const double sna = p->scales ? p->scales[3 * c + 2] : 1.0;
         auto key = std::make_tuple((int)p->celltype[c], skr, scal, sna);
         auto it = psetId.find(key);
         if (it == psetId.end()) {
            CDH_REQUIRE(psets.size() < (size_t)kMaxParamSets,
                        "cdh_sim_create: more than 256 (cell type, scale) combinations");
            it = psetId.emplace(key, (int)psets.size()).first;
            psets.push_back(make_param_set(cts[p->celltype[c]], skr, scal, sna));
         }
         pidIn[c] = (uint8_t)it->second;
      }
      if (psets.empty()) psets.push_back(make_param_set(cts[0], 1.0, 1.0, 1.0));

      std::vector<double> coef = scaled_fit_coefficients(p->fit_l, p->fit_m, p->fit_coef, p->dt);
      track(s.get(), s->coef, coef.size());
      s->coef.upload(coef.data(), coef.size());
      track(s.get(), s->params, psets.size());
      s->params.upload(psets.data(), psets.size());

      const int64_t S = s->stride;
      std::vector<uint8_t> pid(S, 0), fl(S, kFlagGatesFixed);
      std::vector<uint32_t> bi(S, 0);
      std::vector<int64_t> bidx(n);
      std::vector<double> st((size_t)N_STATE * S, 0.0);
      std::vector<uint8_t> tissue(nbox, 0);
      std::vector<double> sig(6 * nbox, 0.0);
      std::vector<double> box0(nbox, 0.0);
      const double c9 = s->K.c9;
      for (int64_t k = 0; k < n; k++) {
         const int64_t c = s->perm[k];
         pid[k] = pidIn[c];
         fl[k] = flIn[c];
         bidx[k] = bidxIn[c];
         bi[k] = (uint32_t)bidxIn[c];
         tissue[bidxIn[c]] = 1;
         std::memcpy(&sig[6 * bidxIn[c]], p->sigma + 6 * c, 6 * sizeof(double));
         const CellType &ct = cts[p->celltype[c]];
         for (int q = 0; q < N_STATE; q++) st[q * S + k] = ct.state0[q];
         st[S_DVK_I * S + k] = ct.state0[S_DVK_I] / c9 + ct.Vm0;
         box0[bidxIn[c]] = ct.Vm0;
      }
      std::vector<double> A0(S, 0.0);
      std::vector<float> Aw(18 * S, 0.0f);
      fgr_weights(g, tissue.data(), sig.data(), n, bidx.data(), A0.data(), Aw.data(), S);
      int64_t off64[19];
      fgr_offsets(g.bx, g.by, g.bz, off64);
      for (int k = 0; k < 19; k++) s->off[k] = (int)off64[k];

      track(s.get(), s->pid, S);
      s->pid.upload(pid.data(), S);
      track(s.get(), s->flags, S);
      s->flags.upload(fl.data(), S);
      track(s.get(), s->bi, S);
      s->bi.upload(bi.data(), S);
      if (s->mode == CDH_MODE_MIXED_FUSED) {
         std::vector<float> s32((size_t)mixed_n32 * S, 0.0f);
         std::vector<double> s64((size_t)2 * S, 0.0);
         s->K0cell.assign(S, 0.0);
         for (int64_t k = 0; k < n; k++) {
            const CellType &ct = cts[p->celltype[s->perm[k]]];
            s->K0cell[k] = ct.state0[S_DVK_I];
            for (int q = 0; q < mixed_n32; q++) {
               const int src = mixed_slot_source(q);
               s32[q * S + k] = src < 0 ? 0.0f : (float)ct.state0[src];
            }
            s64[0 * S + k] = ct.state0[S_CA_I];
            s64[1 * S + k] = ct.state0[S_NA_I];
         }
         track(s.get(), s->st32, s32.size());
         upload_blocked(s->st32, s32, mixed_n32, S);
         track(s.get(), s->st64, s64.size());
         upload_blocked(s->st64, s64, 2, S);
      } else {
         track(s.get(), s->state, (size_t)N_STATE * S);
         upload_blocked(s->state, st, N_STATE, S);
         track(s.get(), s->A0, S);
         s->A0.upload(A0.data(), S);
      }
      track(s.get(), s->Aw, 18 * S);
      upload_blocked(s->Aw, Aw, 18, S);
      const int nbuf = s->mode == CDH_MODE_FP64_UNFUSED ? 1 : 2;
      for (int k = 0; k < nbuf; k++) {
         track(s.get(), s->box[k], nbox);
         s->box[k].upload(box0.data(), nbox);
      }

      // stimuli: strict box in grid tuples (elec/BoxStimulus.cc:18-34)
      s->stim.assign(stim, stim + nstim);
      std::vector<uint8_t> mask(S, 0);
      s->stimListOff.assign(1, 0);
      std::vector<int> lists;
      for (int q = 0; q < nstim; q++) {
         CDH_REQUIRE(stim[q].period > 0 && stim[q].duration > 0, "cdh_sim_create: bad pulse");
         for (int64_t k = 0; k < n; k++) {
            const int64_t gg = p->gid[s->perm[k]];
            const int64_t x = gg % p->nx, y = (gg / p->nx) % p->ny, z = gg / ((int64_t)p->nx * p->ny);
            const cdh_box_stim &b = stim[q];
            if (x > b.xMin && x < b.xMax && y > b.yMin && y < b.yMax && z > b.zMin && z < b.zMax) {
               mask[k] |= (uint8_t)(1u << q);
               lists.push_back((int)k);
            }
         }
         s->stimListOff.push_back((int64_t)lists.size());
      }
      track(s.get(), s->stimMask, S);
      s->stimMask.upload(mask.data(), S);
      track(s.get(), s->stimListDev, lists.size() ? lists.size() : 1);
      s->stimListDev.upload(lists.data(), lists.size());

      if (s->mode == CDH_MODE_FP64_UNFUSED) {
         std::vector<double> vm(S, 0.0);
         for (int64_t k = 0; k < n; k++) vm[k] = box0[bidx[k]];
         track(s.get(), s->Vm, S);
         s->Vm.upload(vm.data(), S);
         track(s.get(), s->dVmR, S);
         s->dVmR.zero();
         track(s.get(), s->dVmD, S);
         s->dVmD.zero();
         track(s.get(), s->iStim, S);
         s->iStim.zero();
      }
      std::vector<double> neg(S, -1.0);
      track(s.get(), s->tact0, S);
      s->tact0.upload(neg.data(), S);
      track(s.get(), s->tact, S);
      s->tact.upload(neg.data(), S);
      track(s.get(), s->trec, S);
      s->trec.upload(neg.data(), S);
      if (p->sensors & 2) {
         track(s.get(), s->apdPhase, S);
         s->apdPhase.zero();
         track(s.get(), s->vmax, S);
         s->vmax.zero();
         track(s.get(), s->t90, S);
         s->t90.upload(neg.data(), S);
      }
      CDH_HIP(hipDeviceSynchronize());
      *out = s.release();
   });
}

int cdh_sim_destroy(cdh_sim *s)
{
   if (s) (void)hipStreamSynchronize(s->stream);
   delete s;
   return CDH_OK;
}

int cdh_sim_run(cdh_sim *s, int nsteps)
{
   return guard([&] {
      CDH_REQUIRE(s && nsteps >= 0, "cdh_sim_run: bad arguments");
      for (int k = 0; k < nsteps; k++) step(s);
   });
}

int cdh_sim_sync(cdh_sim *s)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_sync: null");
      CDH_HIP(hipStreamSynchronize(s->stream));
   });
}

int cdh_sim_run_timed(cdh_sim *s, int nsteps, double *ms)
{
   return guard([&] {
      CDH_REQUIRE(s && ms && nsteps >= 0, "cdh_sim_run_timed: bad arguments");
      CDH_HIP(hipEventRecord(s->ev0, s->stream));
      for (int k = 0; k < nsteps; k++) step(s);
      CDH_HIP(hipEventRecord(s->ev1, s->stream));
      CDH_HIP(hipEventSynchronize(s->ev1));
      float t = 0.0f;
      CDH_HIP(hipEventElapsedTime(&t, s->ev0, s->ev1));
      *ms = t;
   });
}

int cdh_sim_get_vm(cdh_sim *s, double *host)
{
   return guard([&] {
      CDH_REQUIRE(s && (host || s->n == 0), "cdh_sim_get_vm: null argument");
      CDH_HIP(hipStreamSynchronize(s->stream));
      std::vector<double> v = engine_vm(s);
      for (int64_t k = 0; k < s->n; k++) host[s->perm[k]] = v[k];
   });
}

int cdh_sim_get_state(cdh_sim *s, double *host)
{
   return guard([&] {
      CDH_REQUIRE(s && (host || s->n == 0), "cdh_sim_get_state: null argument");
      CDH_HIP(hipStreamSynchronize(s->stream));
      const int64_t n = s->n, S = s->stride;
      std::vector<double> st = state19(s);
      for (int q = 0; q < N_STATE; q++)
         for (int64_t k = 0; k < n; k++) host[q * n + s->perm[k]] = st[q * S + k];
   });
}

int cdh_sim_set_state(cdh_sim *s, const double *state, const double *Vm)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_set_state: null");
      CDH_HIP(hipStreamSynchronize(s->stream));
      const int64_t n = s->n, S = s->stride;
      std::vector<double> vm = engine_vm(s);
      if (Vm)
         for (int64_t k = 0; k < n; k++) vm[k] = Vm[s->perm[k]];
      if (state) {
         std::vector<double> st((size_t)N_STATE * S, 0.0);
         for (int q = 0; q < N_STATE; q++)
            for (int64_t k = 0; k < n; k++) st[q * S + k] = state[q * n + s->perm[k]];
         set_state19(s, st, vm);
      }
      if (Vm) {
         std::vector<uint32_t> bi(n);
         s->bi.download(bi.data(), n);
         if (s->mode != CDH_MODE_FP64_UNFUSED) {
            const BoxGeom &g = s->geom;
            std::vector<double> box((size_t)g.bx * g.by * g.bz, 0.0);
            for (int64_t k = 0; k < n; k++) box[bi[k]] = vm[k];
            s->box[s->cur].upload(box.data(), box.size());
         } else {
            std::vector<double> v(S, 0.0);
            for (int64_t k = 0; k < n; k++) v[k] = vm[k];
            s->Vm.upload(v.data(), S);
         }
      }
   });
}

int cdh_sim_get_activation(cdh_sim *s, double *tact0, double *tact, double *trec)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_get_activation: null");
      CDH_HIP(hipStreamSynchronize(s->stream));
      const int64_t n = s->n;
      std::vector<double> v(n);
      DevBuf<double> *src[3] = {&s->tact0, &s->tact, &s->trec};
      double *dst[3] = {tact0, tact, trec};
      for (int q = 0; q < 3; q++) {
         if (!dst[q]) continue;
         src[q]->download(v.data(), n);
         for (int64_t k = 0; k < n; k++) dst[q][s->perm[k]] = v[k];
      }
   });
}

int cdh_sim_get_t90(cdh_sim *s, double *t90)
{
   return guard([&] {
      CDH_REQUIRE(s && t90, "cdh_sim_get_t90: null argument");
      CDH_REQUIRE(s->p.sensors & 2, "cdh_sim_get_t90: the engine was created without APD90");
      CDH_HIP(hipStreamSynchronize(s->stream));
      std::vector<double> v(s->n);
      s->t90.download(v.data(), s->n);
      for (int64_t k = 0; k < s->n; k++) t90[s->perm[k]] = v[k];
   });
}

double cdh_sim_time(const cdh_sim *s) { return s ? s->time : 0.0; }
int cdh_sim_loop(const cdh_sim *s) { return s ? s->loop : 0; }

int cdh_sim_info(const cdh_sim *s, int64_t *device_bytes, int *bx, int *by, int *bz)
{
   if (!s) return CDH_ERR_ARG;
   if (device_bytes) *device_bytes = s->bytes;
   if (bx) *bx = s->geom.bx;
   if (by) *by = s->geom.by;
   if (bz) *bz = s->geom.bz;
   return CDH_OK;
}

}  // extern "C"
