// batch.hip -- C ABI of the multi-slab engine (cdh_batch_*).
#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "../../include/cardioid_hip.h"
#include "engine_host.hpp"
#include "hip_util.hpp"
#include "slab.hpp"

using namespace cdh;

struct cdh_batch {
   std::unique_ptr<Engine> eng;
   TT06Const K{};
   // per slab: the caller's gids (input order) and grid, for checkpoints
   std::map<int, std::vector<int64_t>> gid;
   std::map<int, std::array<int, 3>> grid;
};

namespace {

SlabInput slab_input(const cdh_slab_desc *d, std::vector<CellType> &cts)
{
   CDH_REQUIRE(d, "slab description is null");
   CDH_REQUIRE(d->ncelltype >= 1 && d->ncelltype <= kMaxCellTypes && d->celltypes,
               "slab: 1..8 cell types required");
   cts.resize(d->ncelltype);
   std::memcpy(cts.data(), d->celltypes, sizeof(CellType) * d->ncelltype);
   SlabInput in;
   in.nx = d->nx;
   in.ny = d->ny;
   in.nz = d->nz;
   in.dx = d->dx;
   in.dy = d->dy;
   in.dz = d->dz;
   in.diffusionScale = d->diffusionScale;
   in.ncell = d->ncell;
   in.gid = d->gid;
   in.celltype = d->celltype;
   in.sigma = d->sigma;
   in.ncelltype = d->ncelltype;
   in.celltypes = cts.data();
   in.scales = d->scales;
   in.threshold = d->threshold;
   in.gate_tail_rule = d->gate_tail_rule;
   in.nstim = d->nstim;
   in.stim = d->stim;
   in.check_ranges = d->check_ranges;
   in.vm_min = d->vm_min;
   in.vm_max = d->vm_max;
   in.setup_on_host = d->setup_on_host;
   in.nremote = d->nremote;
   in.remote_gid = d->remote_gid;
   in.remote_sigma = d->remote_sigma;
   return in;
}

Engine &engine(cdh_batch *b)
{
   CDH_REQUIRE(b && b->eng, "batch is null");
   return *b->eng;
}

// input-order copy of a slab-order array
void scatter(const Engine::Slab &s, const double *src, double *dst, int nq = 1)
{
   for (int q = 0; q < nq; q++)
      for (int64_t k = 0; k < s.n; k++) dst[q * s.n + s.perm[k]] = src[q * s.n + k];
}

}  // namespace

extern "C" {

int cdh_batch_create(const cdh_batch_params *p, cdh_batch **out)
{
   return guard([&] {
      CDH_REQUIRE(p && out, "cdh_batch_create: null argument");
      CDH_REQUIRE(!(p->sensors & 2) || (p->sensors & 1),
                  "cdh_batch_create: APD90 needs the activation sensors (sensors bit 0)");
      auto b = std::make_unique<cdh_batch>();
      tt06_constants(&b->K);
      b->eng = std::make_unique<Engine>(p->mode, p->dt, p->fit_l, p->fit_m, p->fit_coef,
                                        p->cell_capacity, p->box_capacity, p->slab_capacity,
                                        p->param_capacity, p->sensors, p->exact_functions,
                                        p->jhTauSmooth);
      *out = b.release();
   });
}

int cdh_batch_destroy(cdh_batch *b)
{
   delete b;
   return CDH_OK;
}

int cdh_slab_extent(const cdh_slab_desc *d, int64_t *cells, int64_t *box_points, int *param_sets)
{
   return guard([&] {
      std::vector<CellType> cts;
      const SlabInput in = slab_input(d, cts);
      int64_t ntile = 0, nbox = 0;
      slab_extent(in, &ntile, &nbox);
      if (cells) *cells = ntile * kBlk;
      if (box_points) *box_points = nbox;
      if (param_sets) {
         // distinct (cell type, scales) combinations
         std::vector<std::vector<double>> seen;
         for (int64_t c = 0; c < in.ncell; c++) {
            std::vector<double> key = {(double)in.celltype[c],
                                       in.scales ? in.scales[3 * c] : 1.0,
                                       in.scales ? in.scales[3 * c + 1] : 1.0,
                                       in.scales ? in.scales[3 * c + 2] : 1.0};
            bool found = false;
            for (const auto &k : seen) found |= k == key;
            if (!found) seen.push_back(key);
            CDH_REQUIRE(seen.size() <= 256, "slab: more than 256 (cell type, scale) combinations");
         }
         *param_sets = (int)seen.size();
      }
   });
}

int cdh_batch_add(cdh_batch *b, const cdh_slab_desc *d, int *slab)
{
   const int rc = guard([&] {
      CDH_REQUIRE(slab, "cdh_batch_add: null slab output");
      Engine &e = engine(b);
      std::vector<CellType> cts;
      SlabInput in = slab_input(d, cts);
      // configuration A's generic gate loop has no v2 loop bounds
      if (e.exact()) in.gate_tail_rule = 0;
      int64_t ntile = 0, nbox = 0;
      slab_extent(in, &ntile, &nbox);
      if (!e.fits(ntile, nbox, 1)) throw Error(CDH_ERR_FULL, "cdh_batch_add: arena full");
      const SlabHost h = prepare_slab(in, b->K.c9);
      const int id = e.add(h);
      if (id < 0) throw Error(CDH_ERR_FULL, "cdh_batch_add: arena full");
      b->gid[id].assign(d->gid, d->gid + d->ncell);
      b->grid[id] = {d->nx, d->ny, d->nz};
      *slab = id;
   });
   return rc;
}

int cdh_batch_remove(cdh_batch *b, int slab)
{
   return guard([&] {
      engine(b).remove(slab);
      b->gid.erase(slab);
      b->grid.erase(slab);
   });
}

int cdh_batch_write_checkpoint(cdh_batch *b, int slab, const char *dir, int ascii,
                               char *state_path, int path_len)
{
   return guard([&] {
      CDH_REQUIRE(dir, "cdh_batch_write_checkpoint: null directory");
      Engine &e = engine(b);
      const Engine::Slab &s = e.slab(slab);
      const int64_t n = s.n;
      std::vector<double> vm(n), st((size_t)N_STATE * n);
      const std::vector<double> v = e.vm(slab), q = e.state19(slab);
      scatter(s, v.data(), vm.data());
      scatter(s, q.data(), st.data(), N_STATE);
      const auto &g = b->grid.at(slab);
      const int rc = cdh_checkpoint_write(dir, "simulate", s.loop, s.time, g[0], g[1], g[2], n,
                                          b->gid.at(slab).data(), vm.data(), st.data(), ascii,
                                          state_path, path_len);
      if (rc != CDH_OK) throw Error(rc, cdh_last_error());
   });
}

int cdh_batch_read_checkpoint(cdh_batch *b, int slab, const char *state_file)
{
   return guard([&] {
      CDH_REQUIRE(state_file, "cdh_batch_read_checkpoint: null file name");
      Engine &e = engine(b);
      const Engine::Slab &s = e.slab(slab);
      const int64_t n = s.n;
      // input order; fields the file lacks keep their values
      std::vector<double> vm(n), st((size_t)N_STATE * n);
      const std::vector<double> v0 = e.vm(slab), q0 = e.state19(slab);
      scatter(s, v0.data(), vm.data());
      scatter(s, q0.data(), st.data(), N_STATE);
      int loop = 0;
      double time = 0.0;
      const int rc = cdh_checkpoint_read(state_file, n, b->gid.at(slab).data(), vm.data(),
                                         st.data(), &loop, &time);
      if (rc != CDH_OK) throw Error(rc, cdh_last_error());
      std::vector<double> vs(n), qs((size_t)N_STATE * n);
      for (int64_t k = 0; k < n; k++) {
         vs[k] = vm[s.perm[k]];
         for (int q = 0; q < N_STATE; q++) qs[q * n + k] = st[q * n + s.perm[k]];
      }
      e.set_state19(slab, qs, vs);
      e.set_vm(slab, vs);
      e.set_clock(slab, time, loop);
   });
}

int cdh_batch_link_halo(cdh_batch *b, int dst, int src, int64_t *npoints)
{
   return guard([&] {
      const int64_t np = engine(b).link_halo(dst, src);
      if (npoints) *npoints = np;
   });
}

int cdh_batch_run(cdh_batch *b, int nsteps)
{
   return guard([&] { engine(b).run(nsteps); });
}

int cdh_batch_run_timed(cdh_batch *b, int nsteps, double *ms)
{
   return guard([&] {
      CDH_REQUIRE(ms, "cdh_batch_run_timed: null output");
      *ms = engine(b).run_timed(nsteps);
   });
}

int cdh_batch_sync(cdh_batch *b)
{
   return guard([&] { engine(b).sync(); });
}

int cdh_batch_set_graph(cdh_batch *b, int steps)
{
   return guard([&] { engine(b).set_graph_steps(steps); });
}

int cdh_batch_set_launch(cdh_batch *b, int waves, int block, int occupancy, int tile_order)
{
   return guard([&] { engine(b).set_launch(LaunchCfg{waves, block, occupancy, tile_order}); });
}

int cdh_batch_get_launch(cdh_batch *b, int *waves, int *block, int *occupancy, int *tile_order)
{
   return guard([&] {
      const LaunchCfg c = engine(b).launch();
      if (waves) *waves = c.waves;
      if (block) *block = c.block;
      if (occupancy) *occupancy = c.occ;
      if (tile_order) *tile_order = c.xcd;
   });
}

int cdh_batch_autotune(cdh_batch *b, int steps, double *best_ms, double *table, int max_rows,
                       int *nrows)
{
   return guard([&] {
      std::vector<std::pair<LaunchCfg, double>> t;
      const double best = engine(b).autotune(steps, &t);
      if (best_ms) *best_ms = best;
      int k = 0;
      for (; k < (int)t.size() && k < max_rows && table; k++) {
         table[5 * k] = t[k].first.waves;
         table[5 * k + 1] = t[k].first.block;
         table[5 * k + 2] = t[k].first.occ;
         table[5 * k + 3] = t[k].first.xcd;
         table[5 * k + 4] = t[k].second;
      }
      if (nrows) *nrows = (int)t.size();
   });
}

int cdh_batch_get_vm(cdh_batch *b, int slab, double *vm)
{
   return guard([&] {
      CDH_REQUIRE(vm, "cdh_batch_get_vm: null output");
      Engine &e = engine(b);
      const std::vector<double> v = e.vm(slab);
      scatter(e.slab(slab), v.data(), vm);
   });
}

int cdh_batch_get_state(cdh_batch *b, int slab, double *state)
{
   return guard([&] {
      CDH_REQUIRE(state, "cdh_batch_get_state: null output");
      Engine &e = engine(b);
      const std::vector<double> st = e.state19(slab);
      scatter(e.slab(slab), st.data(), state, N_STATE);
   });
}

int cdh_batch_get_activation(cdh_batch *b, int slab, double *tact0, double *tact, double *trec)
{
   return guard([&] {
      Engine &e = engine(b);
      const Engine::Slab &s = e.slab(slab);
      std::vector<double> v[3];
      for (auto &x : v) x.resize(s.n);
      e.activation(slab, v[0].data(), v[1].data(), v[2].data());
      double *dst[3] = {tact0, tact, trec};
      for (int q = 0; q < 3; q++)
         if (dst[q]) scatter(s, v[q].data(), dst[q]);
   });
}

int cdh_batch_get_t90(cdh_batch *b, int slab, double *t90)
{
   return guard([&] {
      CDH_REQUIRE(t90, "cdh_batch_get_t90: null output");
      Engine &e = engine(b);
      const Engine::Slab &s = e.slab(slab);
      std::vector<double> v(s.n);
      e.t90(slab, v.data());
      scatter(s, v.data(), t90);
   });
}

int cdh_batch_sensor_counts(cdh_batch *b, int slab, int64_t *n_tact0, int64_t *n_tact,
                            int64_t *n_trec)
{
   return guard([&] { engine(b).sensor_counts(slab, n_tact0, n_tact, n_trec); });
}

int cdh_batch_get_sensors(cdh_batch *b, int slab, double *rows, double *time, int *loop)
{
   return guard([&] {
      CDH_REQUIRE(rows && time && loop, "cdh_batch_get_sensors: null output");
      Engine &e = engine(b);
      const Engine::Slab &s = e.slab(slab);
      std::vector<double> v((size_t)Engine::kSensorFields * s.n);
      e.sensor_state(slab, v.data(), time, loop);
      scatter(s, v.data(), rows, Engine::kSensorFields);
   });
}

int cdh_batch_set_sensors(cdh_batch *b, int slab, const double *rows, double time, int loop)
{
   return guard([&] {
      CDH_REQUIRE(rows, "cdh_batch_set_sensors: null input");
      CDH_REQUIRE(loop >= 0, "cdh_batch_set_sensors: negative loop");
      Engine &e = engine(b);
      const Engine::Slab &s = e.slab(slab);
      const int64_t n = s.n;
      std::vector<double> v((size_t)Engine::kSensorFields * n);
      for (int q = 0; q < Engine::kSensorFields; q++)
         for (int64_t k = 0; k < n; k++) v[q * n + k] = rows[q * n + s.perm[k]];
      e.set_sensor_state(slab, v.data(), time, loop);
   });
}

int cdh_batch_range_report(cdh_batch *b, int slab, int64_t *count, int *loop, int64_t *cell,
                           double *V, double *dVmR, double *dVmD)
{
   return guard([&] {
      Engine &e = engine(b);
      const RangeRec r = e.range_report(slab);
      if (count) *count = (int64_t)r.count;
      if (!r.taken) return;
      if (loop) *loop = r.loop;
      if (cell) *cell = e.slab(slab).perm[r.cell];
      if (V) *V = r.V;
      if (dVmR) *dVmR = r.dVmR;
      if (dVmD) *dVmD = r.dVmD;
   });
}

int cdh_batch_slab_info(cdh_batch *b, int slab, int64_t *ncell, double *time, int *loop)
{
   return guard([&] {
      const Engine::Slab &s = engine(b).slab(slab);
      if (ncell) *ncell = s.n;
      if (time) *time = s.time;
      if (loop) *loop = s.loop;
   });
}

int cdh_batch_info(cdh_batch *b, int64_t *device_bytes, int64_t *cells_used, int64_t *box_used,
                   int *nslab)
{
   return guard([&] {
      Engine &e = engine(b);
      if (device_bytes) *device_bytes = e.device_bytes();
      if (cells_used) *cells_used = e.cells_used();
      if (box_used) *box_used = e.box_used();
      if (nslab) *nslab = e.nslab();
   });
}

}  // extern "C"
