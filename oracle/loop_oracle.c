/*
 * loop_oracle.c -- CPU restatement of Cardioid's omp simulation loop
 * (elec/simulationLoop.cc:170-241).  TEST INFRASTRUCTURE ONLY.
 *
 * Per step, in the reference's order:
 *   iStim = 0; each box stimulus writes iStim[cells] = pulse(time) while
 *     t0 < time < tf and the pulse is on            (Stimulus.hh:21-27, BoxStimulus.cc:53-64)
 *   reaction: per reaction group (one per cell type, members in ascending
 *     local order, ReactionManager.cc:241-261): non-gates then gates
 *                                                     (TT06Dev_Reaction.cc:393-407)
 *   diffusion: scatter Vm into the padded box, 19-point stencil
 *                                                     (FGRDiffusionOMP.cc:75-131)
 *   Vm += dt*(dVmR + dVmD + iStim)                    (simulationLoop.cc:226-227)
 *   time += dt; ++loop                                (simulationLoop.cc:230-231)
 *   sensors at (time, loop)                           (ActivationTimeSensor.cc:99-111,
 *                                                      ActivationAndRecoverySensor.cc:76-92)
 * APD90 has no counterpart in the reference; its definition is stated in
 * oracle.h (ora_sim_get_t90) and at apd_eval() below.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct ora_sim {
   ora_sim_params p;
   int64_t n;
   int bx, by, bz, x0, y0, z0;
   int64_t off[19];
   int64_t *blockIndex;
   float *A;
   double *A0;
   double *VmBlock;
   double *Vm, *dVmR, *dVmD, *iStim;
   double *state;              /* 19*n SoA */
   ora_fitset fits;            /* dt-scaled */
   /* reaction groups */
   int ngroup;
   int64_t **gidx;             /* member local indices per group */
   int64_t *gcount;
   int64_t *gstride;           /* per-group state row stride: count rounded up to 4 */
   double **gVm, **gdV, **gstate; /* gathered per-group buffers, 64-byte aligned */
   /* reaction kernels other than the oracle's own (ora_sim_set_reaction) */
   ora_react_fn react;
   void *react_ctx;
   int react_threads, pointwise_threads;
   /* stimuli */
   int nstim;
   ora_box_stim *stim;
   int64_t **stimList;
   int64_t *stimCount;
   /* sensors */
   uint8_t *act0;
   double *tact0;
   uint8_t *active;
   double *tactThr, *trecThr;
   uint8_t *apdPhase;
   double *vmax, *t90, *vrest;
   double time;
   int loop;
};

static int64_t tuple_x(const ora_sim *s, int64_t g) { return g % s->p.nx; }
static int64_t tuple_y(const ora_sim *s, int64_t g) { return (g / s->p.nx) % s->p.ny; }
static int64_t tuple_z(const ora_sim *s, int64_t g) { return g / ((int64_t)s->p.nx * s->p.ny); }

ora_sim *ora_sim_create(const ora_sim_params *p, int nstim, const ora_box_stim *stim)
{
   ora_tt06_init();
   ora_sim *s = (ora_sim *)calloc(1, sizeof(ora_sim));
   s->p = *p;
   const int64_t n = p->ncell;
   s->n = n;
   s->fits = *p->fits;
   s->fits.dt_scaled = 0;
   ora_fitset_scale_dt(&s->fits, p->dt);

   /* bounding box padded by one (DiffusionUtils.cc:19-66) */
   int xmin = 1 << 30, ymin = 1 << 30, zmin = 1 << 30, xmax = -1, ymax = -1, zmax = -1;
   for (int64_t c = 0; c < n; c++) {
      int x = (int)tuple_x(s, p->gid[c]), y = (int)tuple_y(s, p->gid[c]), z = (int)tuple_z(s, p->gid[c]);
      xmin = x < xmin ? x : xmin; xmax = x > xmax ? x : xmax;
      ymin = y < ymin ? y : ymin; ymax = y > ymax ? y : ymax;
      zmin = z < zmin ? z : zmin; zmax = z > zmax ? z : zmax;
   }
   if (n == 0) { xmin = ymin = zmin = 0; xmax = ymax = zmax = 0; }
   s->bx = xmax - xmin + 3; s->by = ymax - ymin + 3; s->bz = zmax - zmin + 3;
   s->x0 = xmin - 1; s->y0 = ymin - 1; s->z0 = zmin - 1;
   const size_t nbox = (size_t)s->bx * (size_t)s->by * (size_t)s->bz;
   ora_fgr_offsets(s->bx, s->by, s->bz, s->off);

   s->blockIndex = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
   uint8_t *tissue = (uint8_t *)calloc(nbox, 1);
   double *sigBox = (double *)calloc(6 * nbox, sizeof(double));
   for (int64_t c = 0; c < n; c++) {
      int64_t g = p->gid[c];
      int64_t lx = tuple_x(s, g) - s->x0, ly = tuple_y(s, g) - s->y0, lz = tuple_z(s, g) - s->z0;
      int64_t ib = lz + (int64_t)s->bz * (ly + (int64_t)s->by * lx);
      s->blockIndex[c] = ib;
      tissue[ib] = 1;
      memcpy(sigBox + 6 * ib, p->sigma + 6 * c, 6 * sizeof(double));
   }
   s->A = (float *)malloc(sizeof(float) * 19 * (n ? n : 1));
   s->A0 = (double *)malloc(sizeof(double) * (n ? n : 1));
   ora_fgr_weights(s->bx, s->by, s->bz, tissue, sigBox, p->dx, p->dy, p->dz, n,
                   s->blockIndex, s->A, s->A0);
   free(tissue);
   free(sigBox);
   s->VmBlock = (double *)calloc(nbox, sizeof(double));

   s->Vm = (double *)calloc(n, sizeof(double));
   s->dVmR = (double *)calloc(n, sizeof(double));
   s->dVmD = (double *)calloc(n, sizeof(double));
   s->iStim = (double *)calloc(n, sizeof(double));
   s->state = (double *)calloc(ORA_NSTATE * n, sizeof(double));

   /* reaction groups: one per cell type */
   s->ngroup = p->ncelltype;
   s->gidx = (int64_t **)calloc(s->ngroup, sizeof(int64_t *));
   s->gcount = (int64_t *)calloc(s->ngroup, sizeof(int64_t));
   s->gVm = (double **)calloc(s->ngroup, sizeof(double *));
   s->gdV = (double **)calloc(s->ngroup, sizeof(double *));
   s->gstate = (double **)calloc(s->ngroup, sizeof(double *));
   s->gstride = (int64_t *)calloc(s->ngroup, sizeof(int64_t));
   for (int64_t c = 0; c < n; c++) s->gcount[p->celltype[c]]++;
   for (int g = 0; g < s->ngroup; g++) {
      /* rows of 4-cell multiples from 64-byte aligned bases: the reference's
       * v2 gate kernels load 4 doubles at a time from 32-byte aligned
       * addresses (TT06Dev_Reaction.cc:97, :180-192) */
      int64_t m = s->gcount[g] ? s->gcount[g] : 1;
      int64_t ms = (m + 3) / 4 * 4;
      s->gstride[g] = ms;
      s->gidx[g] = (int64_t *)malloc(sizeof(int64_t) * m);
      if (posix_memalign((void **)&s->gVm[g], 64, sizeof(double) * ms) ||
          posix_memalign((void **)&s->gdV[g], 64, sizeof(double) * ms) ||
          posix_memalign((void **)&s->gstate[g], 64, sizeof(double) * ORA_NSTATE * ms))
         abort();
      memset(s->gVm[g], 0, sizeof(double) * ms);
      memset(s->gdV[g], 0, sizeof(double) * ms);
      memset(s->gstate[g], 0, sizeof(double) * ORA_NSTATE * ms);
      s->gcount[g] = 0;
   }
   s->react_threads = s->pointwise_threads = p->nthreads;
   for (int64_t c = 0; c < n; c++) {
      int g = p->celltype[c];
      s->gidx[g][s->gcount[g]++] = c;
   }
   /* initial state (TT06Dev_Reaction.cc:180-200, :455-460) */
   const double c9 = ora_c9();
   for (int64_t c = 0; c < n; c++) {
      const ora_celltype *ct = &p->celltypes[p->celltype[c]];
      for (int k = 0; k < ORA_NSTATE; k++) s->state[k * n + c] = ct->state0[k];
      s->state[ORA_K_I * n + c] = ct->state0[ORA_K_I] / c9 + ct->Vm0;
      s->Vm[c] = ct->Vm0;
   }

   /* stimuli: cells strictly inside the box (BoxStimulus.cc:18-34) */
   s->nstim = nstim;
   s->stim = (ora_box_stim *)malloc(sizeof(ora_box_stim) * (nstim ? nstim : 1));
   s->stimList = (int64_t **)calloc(nstim ? nstim : 1, sizeof(int64_t *));
   s->stimCount = (int64_t *)calloc(nstim ? nstim : 1, sizeof(int64_t));
   for (int k = 0; k < nstim; k++) {
      s->stim[k] = stim[k];
      s->stimList[k] = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
      for (int64_t c = 0; c < n; c++) {
         int64_t g = p->gid[c];
         int64_t x = tuple_x(s, g), y = tuple_y(s, g), z = tuple_z(s, g);
         if (x > stim[k].xMin && x < stim[k].xMax && y > stim[k].yMin && y < stim[k].yMax &&
             z > stim[k].zMin && z < stim[k].zMax)
            s->stimList[k][s->stimCount[k]++] = c;
      }
   }

   s->act0 = (uint8_t *)calloc(n, 1);
   s->tact0 = (double *)malloc(sizeof(double) * (n ? n : 1));
   s->active = (uint8_t *)calloc(n, 1);
   s->tactThr = (double *)malloc(sizeof(double) * (n ? n : 1));
   s->trecThr = (double *)malloc(sizeof(double) * (n ? n : 1));
   for (int64_t c = 0; c < n; c++) s->tact0[c] = s->tactThr[c] = s->trecThr[c] = -1.0;
   s->apdPhase = (uint8_t *)calloc(n ? n : 1, 1);
   s->vmax = (double *)calloc(n ? n : 1, sizeof(double));
   s->t90 = (double *)malloc(sizeof(double) * (n ? n : 1));
   s->vrest = (double *)malloc(sizeof(double) * (n ? n : 1));
   for (int64_t c = 0; c < n; c++) {
      s->t90[c] = -1.0;
      s->vrest[c] = p->celltypes[p->celltype[c]].Vm0;
   }
   s->time = 0.0;
   s->loop = 0;
   return s;
}

void ora_sim_destroy(ora_sim *s)
{
   if (!s) return;
   for (int g = 0; g < s->ngroup; g++) {
      free(s->gidx[g]); free(s->gVm[g]); free(s->gdV[g]); free(s->gstate[g]);
   }
   free(s->gidx); free(s->gcount); free(s->gVm); free(s->gdV); free(s->gstate);
   free(s->gstride);
   for (int k = 0; k < s->nstim; k++) free(s->stimList[k]);
   free(s->stimList); free(s->stimCount); free(s->stim);
   free(s->blockIndex); free(s->A); free(s->A0); free(s->VmBlock);
   free(s->Vm); free(s->dVmR); free(s->dVmD); free(s->iStim); free(s->state);
   free(s->act0); free(s->tact0); free(s->active); free(s->tactThr); free(s->trecThr);
   free(s->apdPhase); free(s->vmax); free(s->t90); free(s->vrest);
   free(s);
}

/* PeriodicPulse::eval, elec/PeriodicPulse.cc:22-28 */
static double pulse_eval(const ora_box_stim *b, double time)
{
   double t = time - floor(time / b->period) * b->period;
   if (t >= b->phase && t < b->phase + b->duration) return b->amplitude;
   return 0.0;
}

static void reaction(ora_sim *s)
{
   const ora_sim_params *p = &s->p;
   for (int g = 0; g < s->ngroup; g++) {
      const int64_t m = s->gcount[g];
      if (m == 0) continue;
      const int64_t ms = s->gstride[g];
      const ora_celltype *ct = &p->celltypes[g];
      int64_t *ix = s->gidx[g];
      double *gV = s->gVm[g], *gd = s->gdV[g], *gs = s->gstate[g];
      /* gather into the group's own SoA buffer, as a TT06Dev instance holds it */
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
      for (int64_t i = 0; i < m; i++) {
         gV[i] = s->Vm[ix[i]];
         for (int k = 0; k < ORA_NSTATE; k++) gs[k * ms + i] = s->state[k * s->n + ix[i]];
      }
      /* The reference runs these loops on one thread over the whole group.
       * Per-cell results do not depend on the split except through the v2
       * loops' per-group rules: the tail rule (cells from 4*floor(m/4) keep
       * their gates) and the f gate's form (lanes 0..3 of each 12-cell group
       * below 12*floor(m/12)).  Chunks that start at multiples of 12 see
       * both rules exactly as the whole group does. */
      int nthr = s->react_threads > 0 ? s->react_threads : 1;
      int64_t chunk = ((m + nthr - 1) / nthr + 11) / 12 * 12;
#pragma omp parallel for schedule(static) num_threads(nthr)
      for (int t = 0; t < nthr; t++) {
         int64_t b = (int64_t)t * chunk;
         if (b >= m) continue;
         int64_t e = b + chunk < m ? b + chunk : m;
         int cnt = (int)(e - b);
         if (s->react) {
            s->react(s->react_ctx, g, p->dt, cnt, gV + b, gs + b, ms, gd + b);
            continue;
         }
         const ora_scales *sc = NULL;
         if (p->scales) {
            /* per-cell scales: evaluate cell by cell */
            for (int64_t i = b; i < e; i++) {
               sc = &p->scales[ix[i]];
               if (p->exact) ora_tt06_nongate_exact(p->dt, ct, sc, 1, gV + i, gs + i, ms, gd + i);
               else ora_tt06_nongate(&s->fits, p->dt, ct, sc, 1, gV + i, gs + i, ms, gd + i);
            }
         } else {
            if (p->exact) ora_tt06_nongate_exact(p->dt, ct, NULL, cnt, gV + b, gs + b, ms, gd + b);
            else ora_tt06_nongate(&s->fits, p->dt, ct, NULL, cnt, gV + b, gs + b, ms, gd + b);
         }
         if (p->exact) {
            ora_tt06_gates_exact(p->dt, ct->s_switch, p->jhTauSmooth, cnt, gV + b, gs + b, ms);
         } else if (p->no_loop_bounds) {
            ora_tt06_gates_v2_all(&s->fits, ct->s_switch, cnt, gV + b, gs + b, ms);
         } else {
            ora_tt06_gates_v2(&s->fits, ct->s_switch, cnt, gV + b, gs + b, ms);
         }
      }
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
      for (int64_t i = 0; i < m; i++) {
         s->dVmR[ix[i]] = gd[i];
         for (int k = 0; k < ORA_NSTATE; k++) s->state[k * s->n + ix[i]] = gs[k * ms + i];
      }
   }
}

void ora_sim_set_reaction(ora_sim *s, ora_react_fn fn, void *ctx, int react_threads,
                          int pointwise_threads)
{
   s->react = fn;
   s->react_ctx = ctx;
   s->react_threads = react_threads;
   s->pointwise_threads = pointwise_threads;
}

/* APD90 bookkeeping of one cell after an evaluation at `time`.  phase 0:
 * waiting for activation; 1: activated, tracking Vmax; 2: t90 recorded. */
static void apd_eval(uint8_t *phase, double *vmax, double *t90, double vrest, double v,
                     double time, int activated_now)
{
   if (*phase == 0) {
      if (activated_now) { *phase = 1; *vmax = v; }
   } else if (*phase == 1) {
      if (v > *vmax) {
         *vmax = v;
      } else if (v < vrest + 0.1 * (*vmax - vrest)) {
         *t90 = time;
         *phase = 2;
      }
   }
}

void ora_apd_trace(const double *V, int nt, double t0, double dt, double vrest,
                   double threshold, double *tact, double *t90)
{
   uint8_t phase = 0, active = 0;
   double vmax = 0.0;
   *tact = -1.0;
   *t90 = -1.0;
   for (int k = 1; k < nt; k++) {
      const double time = t0 + k * dt;
      int act_now = 0;
      if (active && V[k] < threshold) active = 0;
      if (!active && V[k] > threshold) {
         active = 1;
         if (*tact < 0) { *tact = time; act_now = 1; }
      }
      apd_eval(&phase, &vmax, t90, vrest, V[k], time, act_now);
   }
}

static void sensors(ora_sim *s)
{
   const double thr = s->p.threshold;
#pragma omp parallel for schedule(static) num_threads(s->pointwise_threads)
   for (int64_t c = 0; c < s->n; c++) {
      double v = s->Vm[c];
      if (!s->act0[c] && v > 0) { s->act0[c] = 1; s->tact0[c] = s->time; }
      if (s->active[c] && v < thr) {
         s->active[c] = 0;
         if (s->trecThr[c] < 0 && s->tactThr[c] >= 0) s->trecThr[c] = s->time;
      }
      int act_now = 0;
      if (!s->active[c] && v > thr) {
         s->active[c] = 1;
         if (s->tactThr[c] < 0) { s->tactThr[c] = s->time; act_now = 1; }
      }
      apd_eval(&s->apdPhase[c], &s->vmax[c], &s->t90[c], s->vrest[c], v, s->time, act_now);
   }
}

void ora_sim_run(ora_sim *s, int nsteps)
{
   const int64_t n = s->n;
   const double dt = s->p.dt;
   for (int step = 0; step < nsteps; step++) {
      memset(s->iStim, 0, sizeof(double) * n);
      for (int k = 0; k < s->nstim; k++) {
         const ora_box_stim *b = &s->stim[k];
         if (!(s->time > b->t0 && s->time < b->tf)) continue;
         double v = pulse_eval(b, s->time);
         if (v == 0) continue;
         for (int64_t i = 0; i < s->stimCount[k]; i++) s->iStim[s->stimList[k][i]] = v;
      }
      reaction(s);
#pragma omp parallel for schedule(static) num_threads(s->p.nthreads)
      for (int64_t c = 0; c < n; c++) s->VmBlock[s->blockIndex[c]] = s->Vm[c];
      ora_fgr_calc(n, s->blockIndex, s->A, s->A0, s->off, s->VmBlock, s->p.diffusionScale, s->dVmD);
#pragma omp parallel for schedule(static) num_threads(s->pointwise_threads)
      for (int64_t c = 0; c < n; c++) s->Vm[c] += dt * (s->dVmR[c] + s->dVmD[c] + s->iStim[c]);
      s->time += dt;
      ++s->loop;
      sensors(s);
   }
}

void ora_sim_get_vm(const ora_sim *s, double *out) { memcpy(out, s->Vm, sizeof(double) * s->n); }
void ora_sim_get_state(const ora_sim *s, double *out) { memcpy(out, s->state, sizeof(double) * ORA_NSTATE * s->n); }
void ora_sim_set_state(ora_sim *s, const double *in, const double *Vm)
{
   if (in) memcpy(s->state, in, sizeof(double) * ORA_NSTATE * s->n);
   if (Vm) memcpy(s->Vm, Vm, sizeof(double) * s->n);
}
/* A restart as the reference makes one: readCheckpoint sets the states and
 * Vm (checkpointIO.cc:209-259), the SIMULATE object's loop and time set the
 * clock, and the sensors, created anew from the deck, start fresh. */
void ora_sim_restart(ora_sim *s, const double *in, const double *Vm, int loop, double time)
{
   ora_sim_set_state(s, in, Vm);
   s->loop = loop;
   s->time = time;
   for (int64_t c = 0; c < s->n; c++) {
      s->act0[c] = 0;
      s->tact0[c] = -1.0;
      s->active[c] = 0;
      s->tactThr[c] = -1.0;
      s->trecThr[c] = -1.0;
      s->apdPhase[c] = 0;
      s->vmax[c] = 0.0;
      s->t90[c] = -1.0;
   }
}
void ora_sim_get_activation_vm0(const ora_sim *s, double *out) { memcpy(out, s->tact0, sizeof(double) * s->n); }
void ora_sim_get_activation_thr(const ora_sim *s, double *out) { memcpy(out, s->tactThr, sizeof(double) * s->n); }
void ora_sim_get_recovery_thr(const ora_sim *s, double *out) { memcpy(out, s->trecThr, sizeof(double) * s->n); }
void ora_sim_get_t90(const ora_sim *s, double *out) { memcpy(out, s->t90, sizeof(double) * s->n); }
double ora_sim_time(const ora_sim *s) { return s->time; }
int ora_sim_loop(const ora_sim *s) { return s->loop; }
