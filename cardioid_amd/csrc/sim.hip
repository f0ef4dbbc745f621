// sim.hip -- the monodomain engine of one anatomy: simulationLoop
// (elec/simulationLoop.cc:150-243) with the reaction, diffusion, stimuli
// and activation sensors device resident.
//
// The fused modes are the multi-slab engine holding this one slab (cells in
// box order; the input order is kept in the slab's permutation for the host
// accessors).  The unfused FP64 mode is the reference's own structure, run
// by the C++ mirror of simulationLoop (host/simulate.hpp) over the drop-in
// objects: one TT06Dev reaction per cell type in a ReactionManager, the FGR
// diffusion, box stimuli and activation sensors, all on the device, cells in
// the caller's (local) order.
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
#include "host/simulate.hpp"
#include "slab.hpp"
#include "tt06_f64.hpp"

using namespace cdh;

struct cdh_sim {
   cdh_sim_params p{};
   int mode = CDH_MODE_FP64_FUSED;
   int64_t n = 0;
   std::vector<int64_t> gid;             // input order, for checkpoints
   TT06Const K{};
   double time = 0.0;
   int loop = 0;
   // fused modes
   std::unique_ptr<Engine> eng;
   int slab = -1;
   // unfused mode: the reference's loop over the drop-in objects
   hipStream_t stream = nullptr;
   hipEvent_t ev0 = nullptr, ev1 = nullptr;
   std::unique_ptr<host::Simulate> loopsim;
   std::vector<host::TT06HipReaction *> reactions;   // owned by the ReactionManager
   std::vector<std::vector<int>> groups;             // local cells of each reaction
   int64_t unfusedBytes = 0;
   BoxGeom unfusedGeom{};
   std::vector<double> vrest;                        // APD90's Vrest per cell (unfused)

   ~cdh_sim()
   {
      if (stream) (void)hipStreamSynchronize(stream);
      loopsim.reset();
      if (ev0) (void)hipEventDestroy(ev0);
      if (ev1) (void)hipEventDestroy(ev1);
      if (stream) (void)hipStreamDestroy(stream);
   }
};

namespace {

void step_n(cdh_sim *s, int nsteps)
{
   if (s->eng) s->eng->run(nsteps);
   else if (s->loopsim) host::simulationLoop(*s->loopsim, nsteps);
   for (int k = 0; k < nsteps; k++) s->time += s->p.dt;
   s->loop += nsteps;
}

hipStream_t sim_stream(cdh_sim *s)
{
   if (s->eng) return s->eng->stream();
   return s->stream;
}

void sim_sync(cdh_sim *s)
{
   if (s->eng) s->eng->sync();
   else if (s->stream) CDH_HIP(hipStreamSynchronize(s->stream));
}

const std::vector<int64_t> &sim_perm(cdh_sim *s)
{
   return s->eng->slab(s->slab).perm;
}

// The sensors, fresh: at the start, and anew at a restart, as the
// reference creates them from the deck and checkpoints hold none
void unfused_sensors(cdh_sim *s, host::Simulate &sim)
{
   sim.sensor.clear();
   if (s->p.sensors)
      sim.sensor.push_back(std::make_unique<host::ActivationSensor>(
          s->n, s->p.threshold, (s->p.sensors & 2) ? s->vrest.data() : nullptr, s->stream));
}

// The reference's set-up for this path: reactions per cell type in
// ascending local order (elec/ReactionManager.cc:184-261), the diffusion
// (elec/diffusionFactory.cc:38-68), membrane voltage from the reactions
// (:455-460 of TT06Dev_Reaction.cc), box stimuli with their strict-box cell
// lists (elec/BoxStimulus.cc:18-34), the activation sensors.
void create_unfused(cdh_sim *s, const cdh_sim_params *p, int nstim, const cdh_box_stim *stim)
{
   CDH_REQUIRE(p->gate_tail_rule || p->exact_functions,
               "cdh_sim_create: the unfused mode is the reference's loop; its reactions always "
               "apply the v2 loop bounds (gate_tail_rule = 1)");
   const int64_t n = p->ncell;
   CDH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
   CDH_HIP(hipEventCreate(&s->ev0));
   CDH_HIP(hipEventCreate(&s->ev1));
   auto sim = std::make_unique<host::Simulate>(s->stream);
   sim->dt = p->dt;
   sim->nLocal = n;
   sim->Vm.alloc(n);
   sim->iStim.alloc(n);
   sim->iStim.zero();
   sim->dVmReaction.alloc(n);
   sim->dVmReaction.zero();
   sim->dVmDiffusion.alloc(n);
   sim->dVmDiffusion.zero();
   sim->checkRange = p->check_ranges;
   sim->vMin = p->vm_min;
   sim->vMax = p->vm_max;
   sim->range.alloc(1);
   sim->range.zero();

   // reactions: one per cell type present, its cells in ascending local order
   s->groups.assign(p->ncelltype, {});
   for (int64_t c = 0; c < n; c++) s->groups[p->celltype[c]].push_back((int)c);
   sim->reaction = std::make_unique<host::ReactionManager>();
   std::vector<std::vector<int>> kept;
   for (int t = 0; t < p->ncelltype; t++) {
      const std::vector<int> &g = s->groups[t];
      if (g.empty()) continue;
      cdh_tt06_params tp{};
      tp.dt = p->dt;
      // configuration B (tolerance 1e-4, mod 1, v2 gates) or A (tolerance 0:
      // the exact functions; mod selects the smoothed h and j rates)
      tp.tolerance = p->exact_functions ? 0.0 : 1e-4;
      tp.mod = p->exact_functions ? p->jhTauSmooth : 1;
      tp.fastGate = 1;
      tp.fastNonGate = 0;
      tp.fit_l = p->fit_l;
      tp.fit_m = p->fit_m;
      tp.fit_coef = p->fit_coef;
      tp.celltype = p->celltypes[t];
      tp.s_kr = tp.s_cal = tp.s_na = 1.0;
      if (p->scales) {
         const double *s0 = p->scales + 3 * g[0];
         for (int c : g)
            CDH_REQUIRE(p->scales[3 * c] == s0[0] && p->scales[3 * c + 1] == s0[1] &&
                            p->scales[3 * c + 2] == s0[2],
                        "cdh_sim_create: the unfused mode takes one set of drug scales per cell "
                        "type (one reaction object per type)");
         tp.s_kr = s0[0];
         tp.s_cal = s0[1];
         tp.s_na = s0[2];
      }
      auto r = host::reactionFactory("TT06Dev", tp, (int)g.size(), s->stream);
      s->reactions.push_back(static_cast<host::TT06HipReaction *>(r.get()));
      sim->reaction->addReaction(std::move(r), g);
      kept.push_back(g);
   }
   s->groups = kept;
   sim->reaction->initializeMembraneVoltage(sim->Vm.p);

   // diffusion over the local cells; tissue iff cellType > 0 (FGRUtils.hh:42)
   std::vector<int> tissueType(n, 1);
   cdh_fgr_params fp{};
   fp.nx = p->nx;
   fp.ny = p->ny;
   fp.nz = p->nz;
   fp.dx = p->dx;
   fp.dy = p->dy;
   fp.dz = p->dz;
   fp.diffusionScale = p->diffusionScale;
   fp.nLocal = n;
   fp.nRemote = 0;
   fp.gid = p->gid;
   fp.cellType = tissueType.data();
   fp.sigma = p->sigma;
   sim->diffusion = host::diffusionFactory("FGR", fp, s->stream);
   s->unfusedGeom = bounding_box(p->nx, p->ny, p->nz, p->gid, n, p->dx, p->dy, p->dz);

   // box stimuli: strict box in grid tuples (elec/BoxStimulus.cc:18-34)
   for (int q = 0; q < nstim; q++) {
      CDH_REQUIRE(stim[q].period > 0 && stim[q].duration > 0, "cdh_sim_create: bad pulse");
      const cdh_box_stim &b = stim[q];
      std::vector<int> cells;
      for (int64_t c = 0; c < n; c++) {
         const int64_t gg = p->gid[c];
         const int64_t x = gg % p->nx, y = (gg / p->nx) % p->ny, z = gg / ((int64_t)p->nx * p->ny);
         if (x > b.xMin && x < b.xMax && y > b.yMin && y < b.yMax && z > b.zMin && z < b.zMax)
            cells.push_back((int)c);
      }
      sim->stimulus.push_back(std::make_unique<host::BoxStimulus>(b, cells, s->stream));
   }
   if (p->sensors & 2) {
      s->vrest.resize(n);
      for (int64_t c = 0; c < n; c++) s->vrest[c] = p->celltypes[p->celltype[c]].Vm0;
   }
   unfused_sensors(s, *sim);
   const BoxGeom &g = s->unfusedGeom;
   s->unfusedBytes = n * (int64_t)(N_STATE * 8 + 2 + 4 * 8 + 80 + 4 + 24) +
                     (int64_t)g.bx * g.by * g.bz * 8;
   CDH_HIP(hipDeviceSynchronize());
   s->loopsim = std::move(sim);
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
      if (p->exact_functions && p->mode != CDH_MODE_FP64_UNFUSED &&
          p->mode != CDH_MODE_FP64_FUSED)
         throw Error(kErrUnsupported,
                     "cdh_sim_create: the exact functions (TT06Dev tolerance 0) run in the "
                     "unfused and FP64 fused modes only");
      CDH_REQUIRE(p->exact_functions || (p->fit_l && p->fit_m && p->fit_coef),
                  "cdh_sim_create: null fit arrays");
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
      if (p->ncell > 0) s->gid.assign(p->gid, p->gid + p->ncell);
      tt06_constants(&s->K);
      if (!p->exact_functions) scaled_fit_coefficients(p->fit_l, p->fit_m, p->fit_coef, p->dt);
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
         // configuration A's generic gate loop has no v2 loop bounds
         in.gate_tail_rule = p->exact_functions ? 0 : p->gate_tail_rule;
         in.nstim = nstim;
         in.stim = stim;
         in.check_ranges = p->check_ranges;
         in.vm_min = p->vm_min;
         in.vm_max = p->vm_max;
         in.setup_on_host = p->setup_on_host;
         if (p->mode == CDH_MODE_FP64_UNFUSED) {
            create_unfused(s.get(), p, nstim, stim);
         } else {
            const SlabHost h = prepare_slab(in, s->K.c9);
            s->eng = std::make_unique<Engine>(p->mode, p->dt, p->fit_l, p->fit_m, p->fit_coef,
                                              h.S, h.nbox, 1, (int)h.psets.size(), p->sensors,
                                              p->exact_functions, p->jhTauSmooth);
            LaunchCfg c;
            c.waves = p->variant ? p->variant : 3;
            c.block = 64;
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
      if (!s->loopsim) {
         step_n(s, nsteps);
         *ms = 0.0;
         return;
      }
      CDH_HIP(hipEventRecord(s->ev0, s->stream));
      step_n(s, nsteps);
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
      if (s->n == 0) return;
      sim_sync(s);
      if (s->loopsim) {
         s->loopsim->Vm.download(host, s->n);
         return;
      }
      const std::vector<double> v = s->eng->vm(s->slab);
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
      if (s->loopsim) {
         for (size_t r = 0; r < s->groups.size(); r++) {
            const std::vector<int> &g = s->groups[r];
            std::vector<double> st((size_t)N_STATE * g.size());
            s->reactions[r]->getState(st.data());
            for (int q = 0; q < N_STATE; q++)
               for (size_t k = 0; k < g.size(); k++) host[q * n + g[k]] = st[q * g.size() + k];
         }
         return;
      }
      const std::vector<double> st = s->eng->state19(s->slab);
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
      if (s->loopsim) {
         if (state) {
            for (size_t r = 0; r < s->groups.size(); r++) {
               const std::vector<int> &g = s->groups[r];
               std::vector<double> st((size_t)N_STATE * g.size());
               for (int q = 0; q < N_STATE; q++)
                  for (size_t k = 0; k < g.size(); k++) st[q * g.size() + k] = state[q * n + g[k]];
               s->reactions[r]->setState(st.data());
            }
         }
         if (Vm) s->loopsim->Vm.upload(Vm, n);
         return;
      }
      const std::vector<int64_t> &perm = sim_perm(s);
      std::vector<double> vm = s->eng->vm(s->slab);
      if (Vm)
         for (int64_t k = 0; k < n; k++) vm[k] = Vm[perm[k]];
      if (state) {
         std::vector<double> st((size_t)N_STATE * n);
         for (int q = 0; q < N_STATE; q++)
            for (int64_t k = 0; k < n; k++) st[q * n + k] = state[q * n + perm[k]];
         s->eng->set_state19(s->slab, st, vm);
      }
      if (Vm) s->eng->set_vm(s->slab, vm);
   });
}

int cdh_sim_get_activation(cdh_sim *s, double *tact0, double *tact, double *trec)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_get_activation: null");
      if (s->n == 0) return;
      CDH_REQUIRE(s->p.sensors & 1, "cdh_sim_get_activation: the engine was created without sensors");
      sim_sync(s);
      if (s->loopsim) {
         s->loopsim->sensor[0]->get(tact0, tact, trec);
         return;
      }
      const int64_t n = s->n;
      const std::vector<int64_t> &perm = sim_perm(s);
      std::vector<double> v[3];
      for (auto &x : v) x.resize(n);
      s->eng->activation(s->slab, v[0].data(), v[1].data(), v[2].data());
      double *dst[3] = {tact0, tact, trec};
      for (int q = 0; q < 3; q++)
         if (dst[q])
            for (int64_t k = 0; k < n; k++) dst[q][perm[k]] = v[q][k];
   });
}

int cdh_sim_sensor_counts(cdh_sim *s, int64_t *n_tact0, int64_t *n_tact, int64_t *n_trec)
{
   return guard([&] {
      CDH_REQUIRE(s, "cdh_sim_sensor_counts: null");
      CDH_REQUIRE(s->p.sensors & 1, "cdh_sim_sensor_counts: the engine was created without sensors");
      if (s->n == 0) {
         if (n_tact0) *n_tact0 = 0;
         if (n_tact) *n_tact = 0;
         if (n_trec) *n_trec = 0;
         return;
      }
      sim_sync(s);
      if (s->loopsim) {
         s->loopsim->sensor[0]->counts(n_tact0, n_tact, n_trec);
         return;
      }
      s->eng->sensor_counts(s->slab, n_tact0, n_tact, n_trec);
   });
}

int cdh_sim_get_t90(cdh_sim *s, double *t90)
{
   return guard([&] {
      CDH_REQUIRE(s && t90, "cdh_sim_get_t90: null argument");
      CDH_REQUIRE(s->p.sensors & 2, "cdh_sim_get_t90: the engine was created without APD90");
      if (s->n == 0) return;
      sim_sync(s);
      if (s->loopsim) {
         s->loopsim->sensor[0]->t90(t90);
         return;
      }
      std::vector<double> v(s->n);
      s->eng->t90(s->slab, v.data());
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
      int64_t c = -1;
      if (s->eng) {
         r = s->eng->range_report(s->slab);
         if (r.taken) c = sim_perm(s)[r.cell];
      } else if (s->loopsim) {
         sim_sync(s);
         s->loopsim->range.download(&r, 1);
         s->loopsim->range.zero();
         c = r.cell;
      }
      if (count) *count = (int64_t)r.count;
      if (!r.taken) return;
      if (loop) *loop = r.loop;
      if (cell) *cell = c;
      if (V) *V = r.V;
      if (dVmR) *dVmR = r.dVmR;
      if (dVmD) *dVmD = r.dVmD;
   });
}

int cdh_sim_write_checkpoint(cdh_sim *s, const char *dir, int ascii, char *state_path,
                             int path_len)
{
   return guard([&] {
      CDH_REQUIRE(s && dir, "cdh_sim_write_checkpoint: null argument");
      std::vector<double> vm(s->n), st((size_t)N_STATE * s->n);
      if (s->n) {
         int rc = cdh_sim_get_vm(s, vm.data());
         if (rc == CDH_OK) rc = cdh_sim_get_state(s, st.data());
         if (rc != CDH_OK) throw Error(rc, cdh_last_error());
      }
      const int rc = cdh_checkpoint_write(dir, "simulate", s->loop, s->time, s->p.nx, s->p.ny,
                                          s->p.nz, s->n, s->gid.data(), vm.data(), st.data(),
                                          ascii, state_path, path_len);
      if (rc != CDH_OK) throw Error(rc, cdh_last_error());
   });
}

int cdh_sim_read_checkpoint(cdh_sim *s, const char *state_file)
{
   return guard([&] {
      CDH_REQUIRE(s && state_file, "cdh_sim_read_checkpoint: null argument");
      // fields the file lacks keep their values
      std::vector<double> vm(s->n), st((size_t)N_STATE * s->n);
      if (s->n) {
         int rc = cdh_sim_get_vm(s, vm.data());
         if (rc == CDH_OK) rc = cdh_sim_get_state(s, st.data());
         if (rc != CDH_OK) throw Error(rc, cdh_last_error());
      }
      int loop = 0;
      double time = 0.0;
      int rc = cdh_checkpoint_read(state_file, s->n, s->gid.data(), vm.data(), st.data(), &loop,
                                   &time);
      if (rc != CDH_OK) throw Error(rc, cdh_last_error());
      if (s->n) {
         rc = cdh_sim_set_state(s, st.data(), vm.data());
         if (rc != CDH_OK) throw Error(rc, cdh_last_error());
      }
      s->loop = loop;
      s->time = time;
      if (s->eng) s->eng->set_clock(s->slab, time, loop);
      if (s->loopsim) {
         CDH_HIP(hipStreamSynchronize(s->stream));
         s->loopsim->time = time;
         s->loopsim->loop = loop;
         unfused_sensors(s, *s->loopsim);
      }
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
   } else if (s->loopsim) {
      bytes = s->unfusedBytes;
      g = s->unfusedGeom;
   }
   if (device_bytes) *device_bytes = bytes;
   if (bx) *bx = g.bx;
   if (by) *by = g.by;
   if (bz) *bz = g.bz;
   return CDH_OK;
}

}  // extern "C"
