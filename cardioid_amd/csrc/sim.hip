// sim.hip -- the monodomain engine of one anatomy: simulationLoop
// (elec/simulationLoop.cc:150-243) with the reaction, diffusion, stimuli
// and activation sensors device resident.
//
// The fused modes are the multi-slab engine holding this one slab.  The
// unfused FP64 mode keeps the reference's structure, one kernel per stage:
// iStim = 0 and the box stimuli, TT06Dev_Reaction::calc, the scatter of Vm
// into the box and FGRDiffusionOMP::calc, then the integrator with the
// sensors.  Both hold the cells in box order; the input order is kept in
// the slab's permutation for the host accessors.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/cardioid_hip.h"
#include "engine_host.hpp"
#include "hip_util.hpp"
#include "kernels.hpp"
#include "setup.hpp"
#include "slab.hpp"
#include "tt06_f64.hpp"

namespace cdh {

__global__ void k_integrate_sensor_f64(int n, double dt, double *Vm, const double *dVmR,
                                       const double *dVmD, const double *iStim, SensorArgs s,
                                       const ParamSet *params, const uint8_t *pid);

}  // namespace cdh

using namespace cdh;

struct cdh_sim {
   cdh_sim_params p{};
   int mode = CDH_MODE_FP64_FUSED;
   int64_t n = 0;
   TT06Const K{};
   double time = 0.0;
   int loop = 0;
   // fused modes
   std::unique_ptr<Engine> eng;
   int slab = -1;
   // unfused mode
   struct Unfused {
      int64_t S = 0;
      BoxGeom geom{};
      std::vector<int64_t> perm;
      std::vector<cdh_box_stim> stim;
      std::vector<int64_t> stimListOff;
      hipStream_t stream = nullptr;
      hipEvent_t ev0 = nullptr, ev1 = nullptr;
      int off[19] = {0};
      DevBuf<double> coef;
      DevBuf<ParamSet> params;
      DevBuf<uint8_t> pid, flags;
      DevBuf<double> state;              // 19 fields in 64-cell blocks
      DevBuf<uint32_t> bi;
      DevBuf<double> A0;
      DevBuf<float> Aw;                  // 18 fields in 64-cell blocks
      DevBuf<double> box, Vm, dVmR, dVmD, iStim;
      DevBuf<int> stimListDev;
      DevBuf<double> tact0, tact, trec;
      DevBuf<uint8_t> apdPhase;
      DevBuf<double> vmax, t90;
      DevBuf<RangeRec> range;
      int64_t bytes = 0;
      ~Unfused()
      {
         if (stream) (void)hipStreamSynchronize(stream);
         if (ev0) (void)hipEventDestroy(ev0);
         if (ev1) (void)hipEventDestroy(ev1);
         if (stream) (void)hipStreamDestroy(stream);
      }
   };
   std::unique_ptr<Unfused> u;
};

namespace {

constexpr int kBlock = 256;

template <class T>
void track(cdh_sim::Unfused *u, DevBuf<T> &b, size_t count)
{
   b.alloc(count);
   u->bytes += (int64_t)(count * sizeof(T));
}

template <class T>
std::vector<T> to_blocked(const std::vector<T> &soa, int nq, int64_t S)
{
   std::vector<T> b(soa.size());
   for (int q = 0; q < nq; q++)
      for (int64_t k = 0; k < S; k++) b[blk_index(q, k, nq)] = soa[q * S + k];
   return b;
}

// PeriodicPulse::eval (elec/PeriodicPulse.cc:22-28)
double pulse(const cdh_box_stim &b, double time)
{
   double t = time - std::floor(time / b.period) * b.period;
   if (t >= b.phase && t < b.phase + b.duration) return b.amplitude;
   return 0.0;
}

void step_unfused(cdh_sim *s)
{
   cdh_sim::Unfused *u = s->u.get();
   const int n = (int)s->n;
   const dim3 g(grid_for(n, kBlock)), b(kBlock);
   // iStim = 0; each active stimulus assigns its value on its cells in order
   // (Stimulus::stim, elec/Stimulus.hh:21-27; BoxStimulus.cc:53-64)
   hipLaunchKernelGGL(k_fill_f64, g, b, 0, u->stream, n, 0.0, u->iStim.p);
   for (size_t k = 0; k < u->stim.size(); k++) {
      const cdh_box_stim &st = u->stim[k];
      if (!(s->time > st.t0 && s->time < st.tf)) continue;
      const double v = pulse(st, s->time);
      if (v == 0) continue;
      const int cnt = (int)(u->stimListOff[k + 1] - u->stimListOff[k]);
      if (cnt == 0) continue;
      hipLaunchKernelGGL(k_set_list_f64, dim3(grid_for(cnt, kBlock)), b, 0, u->stream, cnt,
                         u->stimListDev.p + u->stimListOff[k], v, u->iStim.p);
   }
   // reaction
   ReactionArgs r;
   r.K = s->K;
   r.coef = u->coef.p;
   r.params = u->params.p;
   r.pid = u->pid.p;
   r.flags = u->flags.p;
   r.state = u->state.p;
   r.stride = u->S;
   r.n = n;
   r.index = nullptr;
   r.Vm = u->Vm.p;
   r.dVm = u->dVmR.p;
   r.dt = s->p.dt;
   r.gate_mask = 0xFFFu;
   hipLaunchKernelGGL(k_tt06_calc_f64<true>, g, b, 0, u->stream, r);
   // diffusion: scatter Vm into the box, then the stencil
   hipLaunchKernelGGL(k_scatter_f64, g, b, 0, u->stream, n, u->bi.p, u->Vm.p, u->box.p);
   StencilArgs d;
   d.n = n;
   d.stride = u->S;
   d.bi = u->bi.p;
   d.box = u->box.p;
   d.A0 = u->A0.p;
   d.Aw = u->Aw.p;
   for (int k = 0; k < 19; k++) d.off[k] = u->off[k];
   d.scale = s->p.diffusionScale;
   d.dVm = u->dVmD.p;
   hipLaunchKernelGGL(k_fgr_calc_f64<true>, g, b, 0, u->stream, d);
   if (s->p.check_ranges)
      hipLaunchKernelGGL(k_check_range_f64, g, b, 0, u->stream, n, u->Vm.p, u->dVmR.p, u->dVmD.p,
                         s->p.vm_min, s->p.vm_max, s->loop, u->range.p);
   // integrator and sensors at the time after the step
   s->time += s->p.dt;
   ++s->loop;
   SensorArgs a{};
   a.n = n;
   a.enabled = s->p.sensors;
   a.first_eval = s->loop == 1;
   a.time = s->time;
   a.threshold = s->p.threshold;
   a.tact0 = u->tact0.p;
   a.tact = u->tact.p;
   a.trec = u->trec.p;
   a.apd = (s->p.sensors & 2) ? 1 : 0;
   a.apd_phase = u->apdPhase.p;
   a.vmax = u->vmax.p;
   a.t90 = u->t90.p;
   hipLaunchKernelGGL(k_integrate_sensor_f64, g, b, 0, u->stream, n, s->p.dt, u->Vm.p, u->dVmR.p,
                      u->dVmD.p, u->iStim.p, a, u->params.p, u->pid.p);
   CDH_LAUNCH_CHECK();
}

void step_n(cdh_sim *s, int nsteps)
{
   if (s->eng) {
      s->eng->run(nsteps);
      for (int k = 0; k < nsteps; k++) s->time += s->p.dt;
      s->loop += nsteps;
      return;
   }
   for (int k = 0; k < nsteps; k++) {
      if (s->u) {
         step_unfused(s);
      } else {
         s->time += s->p.dt;
         ++s->loop;
      }
   }
}

hipStream_t sim_stream(cdh_sim *s)
{
   if (s->eng) return s->eng->stream();
   if (s->u) return s->u->stream;
   return nullptr;
}

void sim_sync(cdh_sim *s)
{
   if (s->eng) s->eng->sync();
   else if (s->u) CDH_HIP(hipStreamSynchronize(s->u->stream));
}

// slab-order views of the unfused mode's device arrays
std::vector<double> unfused_vm(cdh_sim *s)
{
   std::vector<double> v(s->n);
   if (s->n) s->u->Vm.download(v.data(), s->n);
   return v;
}

std::vector<double> unfused_state(cdh_sim *s)
{
   cdh_sim::Unfused *u = s->u.get();
   const int64_t S = u->S, n = s->n;
   std::vector<double> b((size_t)N_STATE * S);
   u->state.download(b.data(), b.size());
   std::vector<double> st((size_t)N_STATE * n);
   for (int q = 0; q < N_STATE; q++)
      for (int64_t k = 0; k < n; k++) st[q * n + k] = b[blk_index(q, k, N_STATE)];
   return st;
}

const std::vector<int64_t> &sim_perm(cdh_sim *s)
{
   if (s->eng) return s->eng->slab(s->slab).perm;
   return s->u->perm;
}

void create_unfused(cdh_sim *s, const SlabHost &h, const std::vector<double> &coef)
{
   auto u = std::make_unique<cdh_sim::Unfused>();
   const int64_t S = h.S, n = h.n;
   u->S = S;
   u->geom = h.geom;
   u->perm = h.perm;
   u->stim = h.stim;
   for (int k = 0; k < 19; k++) u->off[k] = (int)h.off[k];
   CDH_HIP(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
   CDH_HIP(hipEventCreate(&u->ev0));
   CDH_HIP(hipEventCreate(&u->ev1));
   track(u.get(), u->coef, coef.size());
   u->coef.upload(coef.data(), coef.size());
   track(u.get(), u->params, h.psets.size());
   u->params.upload(h.psets.data(), h.psets.size());
   std::vector<uint8_t> pid(S), fl(S);
   for (int64_t k = 0; k < S; k++) {
      pid[k] = (uint8_t)(h.meta[k] & 0xFFu);
      fl[k] = (uint8_t)((h.meta[k] >> 8) & 0xFFu);
   }
   track(u.get(), u->pid, S);
   u->pid.upload(pid.data(), S);
   track(u.get(), u->flags, S);
   u->flags.upload(fl.data(), S);
   track(u.get(), u->bi, S);
   u->bi.upload(h.bi.data(), S);
   track(u.get(), u->state, (size_t)N_STATE * S);
   const std::vector<double> st = to_blocked(h.state, N_STATE, S);
   u->state.upload(st.data(), st.size());
   track(u.get(), u->A0, S);
   u->A0.upload(h.A0.data(), S);
   track(u.get(), u->Aw, (size_t)18 * S);
   const std::vector<float> aw = to_blocked(h.Aw, 18, S);
   u->Aw.upload(aw.data(), aw.size());
   track(u.get(), u->box, h.nbox);
   u->box.upload(h.box0.data(), h.nbox);
   std::vector<double> vm(S, 0.0);
   for (int64_t k = 0; k < n; k++) vm[k] = h.box0[h.bi[k]];
   track(u.get(), u->Vm, S);
   u->Vm.upload(vm.data(), S);
   track(u.get(), u->dVmR, S);
   u->dVmR.zero();
   track(u.get(), u->dVmD, S);
   u->dVmD.zero();
   track(u.get(), u->iStim, S);
   u->iStim.zero();
   std::vector<int> lists;
   u->stimListOff.assign(1, 0);
   for (const auto &l : h.stimList) {
      lists.insert(lists.end(), l.begin(), l.end());
      u->stimListOff.push_back((int64_t)lists.size());
   }
   track(u.get(), u->stimListDev, lists.size() ? lists.size() : 1);
   u->stimListDev.upload(lists.data(), lists.size());
   std::vector<double> neg(S, -1.0);
   track(u.get(), u->range, 1);
   u->range.zero();
   track(u.get(), u->tact0, S);
   u->tact0.upload(neg.data(), S);
   track(u.get(), u->tact, S);
   u->tact.upload(neg.data(), S);
   track(u.get(), u->trec, S);
   u->trec.upload(neg.data(), S);
   if (s->p.sensors & 2) {
      track(u.get(), u->apdPhase, S);
      u->apdPhase.zero();
      track(u.get(), u->vmax, S);
      u->vmax.zero();
      track(u.get(), u->t90, S);
      u->t90.upload(neg.data(), S);
   }
   CDH_HIP(hipDeviceSynchronize());
   s->u = std::move(u);
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
                      p->mode == CDH_MODE_MIXED_FUSED || p->mode == CDH_MODE_FP64_FMA,
                  "cdh_sim_create: unknown mode");
      CDH_REQUIRE(p->dt > 0, "cdh_sim_create: dt must be positive");
      CDH_REQUIRE(p->fit_l && p->fit_m && p->fit_coef, "cdh_sim_create: null fit arrays");
      CDH_REQUIRE(!(p->sensors & 2) || (p->sensors & 1),
                  "cdh_sim_create: APD90 needs the activation sensors (sensors bit 0)");
      if (p->mode == CDH_MODE_MIXED_FUSED)
         CDH_REQUIRE(p->variant == 0 || (p->variant >= 3 && p->variant <= 4),
                     "cdh_sim_create: mixed variant must be 0 (default) or 3..4");
      else
         CDH_REQUIRE(p->variant == 0 || (p->variant >= 2 && p->variant <= 3),
                     "cdh_sim_create: variant must be 0 (default) or 2..3");
      auto s = std::make_unique<cdh_sim>();
      s->p = *p;
      s->mode = p->mode;
      s->n = p->ncell;
      tt06_constants(&s->K);
      std::vector<double> coef = scaled_fit_coefficients(p->fit_l, p->fit_m, p->fit_coef, p->dt);
      if (p->ncell > 0) {
         std::vector<CellType> cts(p->ncelltype);
         std::memcpy(cts.data(), p->celltypes, sizeof(CellType) * p->ncelltype);
         SlabInput in;
         in.nx = p->nx;
         in.ny = p->ny;
         in.nz = p->nz;
         in.dx = p->dx;
         in.dy = p->dy;
         in.dz = p->dz;
         in.diffusionScale = p->diffusionScale;
         in.ncell = p->ncell;
         in.gid = p->gid;
         in.celltype = p->celltype;
         in.sigma = p->sigma;
         in.ncelltype = p->ncelltype;
         in.celltypes = cts.data();
         in.scales = p->scales;
         in.threshold = p->threshold;
         in.gate_tail_rule = p->gate_tail_rule;
         in.nstim = nstim;
         in.stim = stim;
         in.check_ranges = p->check_ranges;
         in.vm_min = p->vm_min;
         in.vm_max = p->vm_max;
         const SlabHost h = prepare_slab(in, s->K.c9);
         if (p->mode == CDH_MODE_FP64_UNFUSED) {
            create_unfused(s.get(), h, coef);
         } else {
            s->eng = std::make_unique<Engine>(p->mode, p->dt, p->fit_l, p->fit_m, p->fit_coef,
                                              h.S, h.nbox, 1, (int)h.psets.size(), p->sensors);
            LaunchCfg c;
            c.waves = p->variant ? p->variant : 3;
            c.block = 256;
            s->eng->set_launch(c);
            s->slab = s->eng->add(h);
            CDH_REQUIRE(s->slab >= 0, "cdh_sim_create: engine arena too small (internal)");
            s->eng->sync();
         }
      }
      *out = s.release();
   });
}

int cdh_sim_destroy(cdh_sim *s)
{
   if (s) {
      hipStream_t st = sim_stream(s);
      if (st) (void)hipStreamSynchronize(st);
   }
   delete s;
   return CDH_OK;
}

int cdh_sim_run(cdh_sim *s, int nsteps)
{
   return guard([&] {
      CDH_REQUIRE(s && nsteps >= 0, "cdh_sim_run: bad arguments");
      step_n(s, nsteps);
   });
}

int cdh_sim_sync(cdh_sim *s)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_sync: null");
      sim_sync(s);
   });
}

int cdh_sim_run_timed(cdh_sim *s, int nsteps, double *ms)
{
   return guard([&] {
      CDH_REQUIRE(s && ms && nsteps >= 0, "cdh_sim_run_timed: bad arguments");
      if (s->eng) {
         *ms = s->eng->run_timed(nsteps);
         for (int k = 0; k < nsteps; k++) s->time += s->p.dt;
         s->loop += nsteps;
         return;
      }
      if (!s->u) {
         step_n(s, nsteps);
         *ms = 0.0;
         return;
      }
      CDH_HIP(hipEventRecord(s->u->ev0, s->u->stream));
      step_n(s, nsteps);
      CDH_HIP(hipEventRecord(s->u->ev1, s->u->stream));
      CDH_HIP(hipEventSynchronize(s->u->ev1));
      float t = 0.0f;
      CDH_HIP(hipEventElapsedTime(&t, s->u->ev0, s->u->ev1));
      *ms = t;
   });
}

int cdh_sim_get_vm(cdh_sim *s, double *host)
{
   return guard([&] {
      CDH_REQUIRE(s && (host || s->n == 0), "cdh_sim_get_vm: null argument");
      if (s->n == 0) return;
      sim_sync(s);
      const std::vector<double> v = s->eng ? s->eng->vm(s->slab) : unfused_vm(s);
      const std::vector<int64_t> &perm = sim_perm(s);
      for (int64_t k = 0; k < s->n; k++) host[perm[k]] = v[k];
   });
}

int cdh_sim_get_state(cdh_sim *s, double *host)
{
   return guard([&] {
      CDH_REQUIRE(s && (host || s->n == 0), "cdh_sim_get_state: null argument");
      if (s->n == 0) return;
      sim_sync(s);
      const int64_t n = s->n;
      const std::vector<double> st = s->eng ? s->eng->state19(s->slab) : unfused_state(s);
      const std::vector<int64_t> &perm = sim_perm(s);
      for (int q = 0; q < N_STATE; q++)
         for (int64_t k = 0; k < n; k++) host[q * n + perm[k]] = st[q * n + k];
   });
}

int cdh_sim_set_state(cdh_sim *s, const double *state, const double *Vm)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_set_state: null");
      if (s->n == 0) return;
      sim_sync(s);
      const int64_t n = s->n;
      const std::vector<int64_t> &perm = sim_perm(s);
      std::vector<double> vm = s->eng ? s->eng->vm(s->slab) : unfused_vm(s);
      if (Vm)
         for (int64_t k = 0; k < n; k++) vm[k] = Vm[perm[k]];
      std::vector<double> st;
      if (state) {
         st.resize((size_t)N_STATE * n);
         for (int q = 0; q < N_STATE; q++)
            for (int64_t k = 0; k < n; k++) st[q * n + k] = state[q * n + perm[k]];
      }
      if (s->eng) {
         if (state) s->eng->set_state19(s->slab, st, vm);
         if (Vm) s->eng->set_vm(s->slab, vm);
         return;
      }
      cdh_sim::Unfused *u = s->u.get();
      if (state) {
         std::vector<double> b((size_t)N_STATE * u->S);
         u->state.download(b.data(), b.size());
         for (int q = 0; q < N_STATE; q++)
            for (int64_t k = 0; k < n; k++) b[blk_index(q, k, N_STATE)] = st[q * n + k];
         u->state.upload(b.data(), b.size());
      }
      if (Vm) u->Vm.upload(vm.data(), n);
   });
}

int cdh_sim_get_activation(cdh_sim *s, double *tact0, double *tact, double *trec)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_get_activation: null");
      if (s->n == 0) return;
      CDH_REQUIRE(s->p.sensors & 1, "cdh_sim_get_activation: the engine was created without sensors");
      sim_sync(s);
      const int64_t n = s->n;
      const std::vector<int64_t> &perm = sim_perm(s);
      std::vector<double> v[3];
      for (auto &x : v) x.resize(n);
      if (s->eng) {
         s->eng->activation(s->slab, v[0].data(), v[1].data(), v[2].data());
      } else {
         s->u->tact0.download(v[0].data(), n);
         s->u->tact.download(v[1].data(), n);
         s->u->trec.download(v[2].data(), n);
      }
      double *dst[3] = {tact0, tact, trec};
      for (int q = 0; q < 3; q++)
         if (dst[q])
            for (int64_t k = 0; k < n; k++) dst[q][perm[k]] = v[q][k];
   });
}

int cdh_sim_get_t90(cdh_sim *s, double *t90)
{
   return guard([&] {
      CDH_REQUIRE(s && t90, "cdh_sim_get_t90: null argument");
      CDH_REQUIRE(s->p.sensors & 2, "cdh_sim_get_t90: the engine was created without APD90");
      if (s->n == 0) return;
      sim_sync(s);
      std::vector<double> v(s->n);
      if (s->eng) s->eng->t90(s->slab, v.data());
      else s->u->t90.download(v.data(), s->n);
      const std::vector<int64_t> &perm = sim_perm(s);
      for (int64_t k = 0; k < s->n; k++) t90[perm[k]] = v[k];
   });
}

int cdh_sim_range_report(cdh_sim *s, int64_t *count, int *loop, int64_t *cell, double *V,
                         double *dVmR, double *dVmD)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_range_report: null");
      RangeRec r{};
      if (s->eng) {
         r = s->eng->range_report(s->slab);
      } else if (s->u) {
         sim_sync(s);
         s->u->range.download(&r, 1);
         s->u->range.zero();
      }
      if (count) *count = (int64_t)r.count;
      if (!r.taken) return;
      if (loop) *loop = r.loop;
      if (cell) *cell = sim_perm(s)[r.cell];
      if (V) *V = r.V;
      if (dVmR) *dVmR = r.dVmR;
      if (dVmD) *dVmD = r.dVmD;
   });
}

double cdh_sim_time(const cdh_sim *s) { return s ? s->time : 0.0; }
int cdh_sim_loop(const cdh_sim *s) { return s ? s->loop : 0; }

int cdh_sim_info(const cdh_sim *s, int64_t *device_bytes, int *bx, int *by, int *bz)
{
   if (!s) return CDH_ERR_ARG;
   int64_t bytes = 0;
   BoxGeom g{};
   if (s->eng) {
      bytes = s->eng->device_bytes();
      g = s->eng->slab(s->slab).geom;
   } else if (s->u) {
      bytes = s->u->bytes;
      g = s->u->geom;
   }
   if (device_bytes) *device_bytes = bytes;
   if (bx) *bx = g.bx;
   if (by) *by = g.by;
   if (bz) *bz = g.bz;
   return CDH_OK;
}

}  // extern "C"
