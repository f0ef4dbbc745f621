/*
 * cardioid_hip.h -- C ABI of the MI355X (gfx950) implementation of
 * Cardioid's electrics hot path: the TT06Dev reaction and the FGR 19-point
 * diffusion stencil, plus the loop pieces that drive them.
 *
 * All pointers named as device arrays (Vm, dVm, iStim, index, ...) are HIP
 * device pointers, i.e. what lazy_array::useOn(GPU) hands out in the
 * reference (elec/lazy_array.hh:204-418).  `stream` is a hipStream_t passed
 * as void* (NULL = the default stream).  Host arrays are plain host memory.
 *
 * Every function returns 0 on success or a CDH_ERR_* code; the message of
 * the last failure on the calling thread is cdh_last_error().  Where the
 * reference asserts on bad input (reactionFactory.cc:70-78,
 * diffusionFactory.cc:65-67) these return CDH_ERR_ARG instead.
 *
 * Reference interfaces replaced, by group:
 *   cdh_tt06_*  -> class Reaction (elec/Reaction.hh:10-43) as implemented by
 *                  TT06Dev_Reaction (elec/TT06Dev_Reaction.cc:83-664) and made
 *                  by REACTION_FACTORY(TT06Dev) (:583-660)
 *   cdh_fgr_*   -> class Diffusion (elec/Diffusion.hh:19-39) as implemented by
 *                  FGRDiffusionOMP (elec/FGRDiffusionOMP.cc:14-265) and chosen
 *                  in diffusionFactory (elec/diffusionFactory.cc:38-103)
 *   cdh_fill / cdh_set_list / cdh_integrate -> the DEVICE_PARALLEL_FORALL
 *                  bodies of simulationLoop (elec/simulationLoop.cc:186-193,
 *                  :221-227) and BoxStimulus::subClassStim (elec/BoxStimulus.cc:53-64)
 *   cdh_activation_sensor_* -> ActivationTimeSensor::eval
 *                  (elec/ActivationTimeSensor.cc:99-111) and
 *                  ActivationAndRecoverySensor::eval (elec/ActivationAndRecoverySensor.cc:76-92)
 *   cdh_sim_*   -> simulationLoop (elec/simulationLoop.cc:150-243) with the
 *                  reaction, diffusion, stimuli and sensors above, device resident
 */
#ifndef CARDIOID_HIP_H
#define CARDIOID_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CDH_OK 0
#define CDH_ERR_ARG 1
#define CDH_ERR_HIP 2
#define CDH_ERR_UNSUPPORTED 3
#define CDH_ERR_STATE 4
#define CDH_ERR_FULL 5         /* the engine's arena has no room for the slab */

#define CDH_NSTATE 19
#define CDH_NFIT 33

const char *cdh_last_error(void);
int cdh_version(void);

/* ------------------------------------------------------------------------ */
/* Cell types (CellTypeParmsFull, elec/TT06Func.hh:49-56).  state0 is in the */
/* TT06_DEV_STATE_INDEX order (elec/TT06Func.h:10-14) with K_i (not dVK_i)   */
/* in slot 1, as getStandardCellTypes fills it.                              */
/* ------------------------------------------------------------------------ */
typedef struct {
   int s_switch;
   double P_NaK, g_Ks, g_Kr, g_to, g_NaL;
   double Vm0;
   double state0[CDH_NSTATE];
} cdh_celltype;

/* TT06Func::getStandardCellTypes (elec/TT06Func.cc:421-612): endoCellML,
 * midCellML, epiCellML, endoRRG, midRRG, epiRRG. */
int cdh_celltype_standard(const char *name, cdh_celltype *out);

/* ------------------------------------------------------------------------ */
/* Reaction: TT06Dev                                                         */
/* ------------------------------------------------------------------------ */
typedef struct cdh_tt06 cdh_tt06;

/* Two configurations of the factory's keywords are implemented:
 *   B: tolerance = 1e-4, mod = 1, fastGate >= 1, fastNonGate = 0 -- the
 *      rational fits of the fit file, the v2 gate kernels, the generic
 *      non-gate update;
 *   A: tolerance <= 0 (the factory's default, TT06Dev_Reaction.cc:600-602)
 *      -- the exact TT06 functions with the generic gate and non-gate
 *      updates (TT06Func.cc:64-161, TT06NonGates.c:319-364); mod = 1 takes
 *      the smoothed h and j rates (TT06Tau.cc:131-151); fastGate and
 *      fastNonGate are ignored, as the reference ignores them there
 *      (:110-114); the fit arrays are not read and may be NULL.
 * Anything else returns CDH_ERR_UNSUPPORTED (e.g. fits with fastGate = 0,
 * which applies dt twice in the reference). */
typedef struct {
   double dt;                  /* time step the fits are scaled by (ms) */
   double tolerance;           /* REACTION keyword: 1e-4 (B) or <= 0 (A) */
   int mod;                    /* REACTION keyword (jhTauSmooth): 1 in B; 0 or 1 in A */
   int fastGate;               /* B: >= 1, the v2 gate kernels */
   int fastNonGate;            /* B: 0, the generic non-gate update */
   const int *fit_l;           /* 33 fits in the order of the fitFile's      */
   const int *fit_m;           /* `functions` list; coefficients packed, l+m */
   const double *fit_coef;     /* per fit (elec/pade.hh:4-26 layout)          */
   cdh_celltype celltype;      /* from cellTypeName / CELLTYPE clone          */
   double s_kr, s_cal, s_na;   /* conductance scales on g_Kr, g_CaL, g_Na (1 = reference) */
} cdh_tt06_params;

/* REACTION_FACTORY(TT06Dev) + TT06Dev_Reaction ctor
 * (elec/TT06Dev_Reaction.cc:583-660, :83-157).  State is allocated on the
 * device and initialised to the cell type's rest state (:180-200). */
int cdh_tt06_create(const cdh_tt06_params *p, int numPoints, cdh_tt06 **out);
int cdh_tt06_destroy(cdh_tt06 *r);

/* Reaction::calc (elec/Reaction.hh:15-17; TT06Dev :393-407): for i < n,
 * reads Vm[indexArray[i]], writes dVm[indexArray[i]] = -Iion, advances the
 * 19 states by dt.  indexArray may be NULL (identity).  iStim is ignored,
 * as TT06Dev ignores it.  Vm and dVm are separate arrays (read-only and
 * write-only in the reference's signature) and must not overlap. */
int cdh_tt06_calc(cdh_tt06 *r, double dt, const int *indexArray, int n, const double *Vm,
                  const double *iStim, double *dVm, void *stream);
/* Reaction::updateNonGate / updateGate (pdr loop; TT06Dev :408-453) */
int cdh_tt06_update_nongate(cdh_tt06 *r, double dt, const int *indexArray, int n,
                            const double *Vm, double *dVR, void *stream);
int cdh_tt06_update_gate(cdh_tt06 *r, double dt, const int *indexArray, int n,
                         const double *Vm, void *stream);
/* Reaction::initializeMembraneVoltage (TT06Dev :455-460) */
int cdh_tt06_initialize_membrane_voltage(cdh_tt06 *r, const int *indexArray, int n,
                                         double *Vm, void *stream);
/* checkpoint accessors: getCheckpointInfo / getVarHandle / getValue / setValue
 * (TT06Dev :462-555).  Field i in 0..18 -> name, unit, handle. */
int cdh_tt06_checkpoint_field(int i, const char **name, const char **unit, int *handle);
int cdh_tt06_var_handle(const char *name);      /* -1 if unknown */
int cdh_tt06_get_value(cdh_tt06 *r, int iCell, int handle, double *value);
int cdh_tt06_set_value(cdh_tt06 *r, int iCell, int handle, double value);
/* bulk state transfer, host array [19][numPoints] */
int cdh_tt06_get_state(cdh_tt06 *r, double *host);
int cdh_tt06_set_state(cdh_tt06 *r, const double *host);

/* ------------------------------------------------------------------------ */
/* Diffusion: FGR                                                            */
/* ------------------------------------------------------------------------ */
typedef struct cdh_fgr cdh_fgr;

typedef struct {
   int nx, ny, nz;             /* anatomy grid; gid = x + nx*(y + ny*z) */
   double dx, dy, dz;          /* mm */
   double diffusionScale;      /* mm^3/mF */
   int64_t nLocal, nRemote;
   const int64_t *gid;         /* nLocal + nRemote cells, local first */
   const int *cellType;        /* tissue iff cellType > 0 (elec/FGRUtils.hh:42) */
   const double *sigma;        /* 6 per cell: a11 a12 a13 a22 a23 a33 (S/mm) */
} cdh_fgr_params;

/* FGRDiffusionOMP ctor: bounding box of the local cells padded by one
 * (elec/DiffusionUtils.cc:19-66), weights (elec/FGRDiffusionOMP.cc:166-265)
 * in float with a double centre weight, as the CMake build has them. */
int cdh_fgr_create(const cdh_fgr_params *p, cdh_fgr **out);
int cdh_fgr_destroy(cdh_fgr *d);
/* Diffusion::updateLocalVoltage / updateRemoteVoltage: scatter into the box */
int cdh_fgr_update_local_voltage(cdh_fgr *d, const double *VmLocal, void *stream);
int cdh_fgr_update_remote_voltage(cdh_fgr *d, const double *VmRemote, void *stream);
/* Diffusion::calc, omp-loop contract: dVm[i] = scale*(stencil) for i < nLocal */
int cdh_fgr_calc(cdh_fgr *d, double *dVm, void *stream);
double cdh_fgr_diffusion_scale(const cdh_fgr *d);

/* ------------------------------------------------------------------------ */
/* Set-up on the device (all pointers device memory, asynchronous on stream  */
/* unless stated)                                                            */
/* ------------------------------------------------------------------------ */
/* FibreConductivity::compute, calcConductivityMatrixIBT
 * (elec/FibreConductivity.cc:21-62): sigma[6c..] = {a11 a12 a13 a22 a23 a33}
 * of fibre angles theta[c], phi[c] (radians), with the device's sin/cos. */
int cdh_fibre_conductivity(int64_t n, const double *theta, const double *phi, double sigmaL,
                           double sigmaT, double *sigma, void *stream);
/* The same for integer angles (angle = i*pi/256, the pio anatomy's integer
 * fields, elec/setConductivity.cc:184-190), |i| <= CDH_FIBRE_ANGLE_MAX: the
 * sines and cosines are the host's, so the tensors equal a host
 * computation bit for bit.  Synchronous; an angle out of range is an error. */
#define CDH_FIBRE_ANGLE_MAX 512
int cdh_fibre_conductivity_int(int64_t n, const int *theta, const int *phi, double sigmaL,
                               double sigmaT, double *sigma, void *stream);
/* cdh_fibre_conductivity_int without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory (a serving
 * loop admitting requests, where hipFree would stall the device).  The
 * angle table (16,400 B) and the out-of-range counter live in the library's
 * own per-device scratch, which holds nothing else and is independent of
 * n; the table is written at the first call on each device.  Calls that
 * overlap on one device race on the counter.  Same arguments, results
 * (bit for bit) and errors as cdh_fibre_conductivity_int; n <= 2^31 - 1. */
int cdh_fibre_conductivity_int_noalloc(int64_t n, const int *theta, const int *phi,
                                       double sigmaL, double sigmaT, double *sigma, void *stream);
/* FGRDiffusionOMP::precomputeCoefficients (elec/FGRDiffusionOMP.cc:166-224)
 * on the device: for ncell cells at box points blockIndex[c] of a
 * bx*by*bz box (z fastest) holding tissue flags and sigma[6*], A0[c]
 * (double) and Aw[(k-1)*stride + c] (float, k = 1..18), rounded as the
 * host build (float weights, one rounding per term, faces in order). */
int cdh_fgr_weights(int bx, int by, int bz, double dx, double dy, double dz,
                    const uint8_t *tissue, const double *sigma, int64_t ncell,
                    const int64_t *blockIndex, double *A0, float *Aw, int64_t stride,
                    void *stream);

/* ------------------------------------------------------------------------ */
/* Loop pieces                                                               */
/* ------------------------------------------------------------------------ */
int cdh_fill(int64_t n, double value, double *x, void *stream);
int cdh_set_list(int64_t n, const int *list, double value, double *x, void *stream);
int cdh_integrate(int64_t n, double dt, double *Vm, const double *dVmR, const double *dVmD,
                  const double *iStim, void *stream);

typedef struct cdh_sensor cdh_sensor;
/* activation bookkeeping for n cells: first Vm > 0 (ActivationTimeSensor),
 * first up-crossing of `threshold` and the first down-crossing after it
 * (ActivationAndRecoverySensor) */
int cdh_activation_sensor_create(int64_t n, double threshold, cdh_sensor **out);
int cdh_activation_sensor_eval(cdh_sensor *s, double time, const double *Vm, void *stream);
int cdh_activation_sensor_get(cdh_sensor *s, double *tact0, double *tact, double *trec);
/* How many cells have their tact0, tact and trec recorded, counted on the
 * device as they are recorded (one wavefront ballot and atomic per kind):
 * the "cells activated" figure without copying the arrays.  No reference
 * counterpart (the reference's sensors keep only the per-cell arrays). */
int cdh_activation_sensor_counts(cdh_sensor *s, int64_t *n_tact0, int64_t *n_tact,
                                 int64_t *n_trec);
int cdh_activation_sensor_destroy(cdh_sensor *s);
/* APD90 on the same sensor (no counterpart in the reference): after the
 * first up-crossing of the threshold, t90 is the first evaluation time with
 * Vm < vrest + 0.1*(Vmax - vrest), Vmax the largest Vm since that crossing.
 * vrest: host array of n rest potentials.  Enable before the first eval. */
int cdh_activation_sensor_enable_apd(cdh_sensor *s, const double *vrest);
int cdh_activation_sensor_get_t90(cdh_sensor *s, double *t90);

/* ------------------------------------------------------------------------ */
/* Monodomain engine: simulationLoop with everything device resident         */
/* ------------------------------------------------------------------------ */
typedef struct cdh_sim cdh_sim;

/* box stimulus with a periodic pulse (stimulusFactory.cc:24-67,
 * BoxStimulus.cc:18-34, PeriodicPulse.cc:22-28): cells with xMin < x < xMax
 * (and y, z likewise) get iStim = amplitude while t0 < time < tf and
 * phase <= time mod period < phase + duration.  amplitude = -vStim. */
typedef struct {
   int xMin, xMax, yMin, yMax, zMin, zMax;
   double t0, tf, period, duration, amplitude, phase;
} cdh_box_stim;

#define CDH_MODE_FP64_UNFUSED 0   /* reference-structured: reaction, scatter, stencil, integrate */
#define CDH_MODE_FP64_FUSED 1     /* one launch per step, same arithmetic */
#define CDH_MODE_MIXED_FUSED 2    /* one launch per step; FP32 gates/concentrations and stencil,
                                     FP64 Vm, Ca_i, Na_i */
#define CDH_MODE_FP64_FMA 3       /* one launch per step, FP64 throughout; the fit polynomials
                                     use fused multiply-adds (rounding differs from the
                                     reference's by ulps; Vm parity within 1e-9) */

typedef struct {
   int nx, ny, nz;
   double dx, dy, dz, dt, diffusionScale;
   int64_t ncell;               /* tissue cells */
   const int64_t *gid;          /* ncell */
   const uint8_t *celltype;     /* ncell, index into celltypes */
   const double *sigma;         /* 6*ncell */
   int ncelltype;               /* <= 8 */
   const cdh_celltype *celltypes;
   const double *scales;        /* 3*ncell (s_kr, s_cal, s_na) or NULL */
   const int *fit_l;
   const int *fit_m;
   const double *fit_coef;
   double threshold;            /* activation threshold, mV (-40 default) */
   int mode;                    /* CDH_MODE_* */
   int gate_tail_rule;          /* 1: the v2 loops' per-group bounds (reference) */
   int sensors;                 /* bit 0: activation times; bit 1: APD90 as well */
   int variant;                 /* fused kernel register target, waves/SIMD (FP64: 2..3,
                                   mixed: 3..4); 0 = default */
   int check_ranges;            /* Simulate::checkRanges (checkRanges, default on) */
   double vm_min, vm_max;       /* its range (VmMin, VmMax: -150, 100 by default,
                                   initializeSimulate.cc:96-98) */
   int setup_on_host;           /* 1: FGR weights built on the host; 0 (default): on
                                   the device, bit for bit the same */
   int exact_functions;         /* 1: TT06Dev configuration A (tolerance 0: the exact
                                   functions, generic gate update; fit arrays may be
                                   NULL); CDH_MODE_FP64_UNFUSED and CDH_MODE_FP64_FUSED
                                   (bit for bit the same step) */
   int jhTauSmooth;             /* with exact_functions: the mod keyword (1: smoothed
                                   h and j rates) */
} cdh_sim_params;

int cdh_sim_create(const cdh_sim_params *p, int nstim, const cdh_box_stim *stim, cdh_sim **out);
int cdh_sim_destroy(cdh_sim *s);
/* advances nsteps omp-loop steps; asynchronous on the engine's stream */
int cdh_sim_run(cdh_sim *s, int nsteps);
int cdh_sim_sync(cdh_sim *s);
/* runs nsteps and returns the GPU time between events recorded on the
 * engine's stream around them, in ms */
int cdh_sim_run_timed(cdh_sim *s, int nsteps, double *ms);
int cdh_sim_get_vm(cdh_sim *s, double *host);          /* ncell, input order */
int cdh_sim_get_state(cdh_sim *s, double *host);       /* [19][ncell], input order */
int cdh_sim_set_state(cdh_sim *s, const double *state, const double *Vm);
int cdh_sim_get_activation(cdh_sim *s, double *tact0, double *tact, double *trec);
/* APD90 end time per cell (-1 if not reached); APD90 = t90 - tact.  No
 * counterpart in the reference: after the first up-crossing of the
 * threshold, t90 is the first time Vm < Vrest + 0.1*(Vmax - Vrest), with
 * Vrest the cell type's rest Vm and Vmax the largest Vm since activation. */
int cdh_sim_get_t90(cdh_sim *s, double *t90);
/* cdh_activation_sensor_counts for the simulation's sensors (any mode) */
int cdh_sim_sensor_counts(cdh_sim *s, int64_t *n_tact0, int64_t *n_tact, int64_t *n_trec);
/* Vm range check (Simulate::checkRanges, Simulate.cc:35-53, before the
 * integrator): the number of (cell, step) pairs whose Vm left [vm_min,
 * vm_max] since the last call, and one of them (the first step's) with the
 * values Simulate::outOfRange prints (loop, cell in input order, V, dVmR,
 * dVmD).  count 0 leaves the other outputs untouched; the call clears the
 * record. */
int cdh_sim_range_report(cdh_sim *s, int64_t *count, int *loop, int64_t *cell, double *V,
                         double *dVmR, double *dVmD);
double cdh_sim_time(const cdh_sim *s);
int cdh_sim_loop(const cdh_sim *s);
/* bytes of device memory held, and the box dimensions used */
int cdh_sim_info(const cdh_sim *s, int64_t *device_bytes, int *bx, int *by, int *bz);

/* ------------------------------------------------------------------------ */
/* Checkpoint / restart files (elec/checkpointIO.cc:121-259)                 */
/* ------------------------------------------------------------------------ */
/* writeCheckpoint: <dir>/snapshot.<loop, 12 digits>/state#000000 (a pio
 * file: `state FILEHEADER` object, blank line, one fixed-length record per
 * cell: gid, Vm, the 19 TT06Dev fields by name) and .../restart (the
 * SIMULATE object's loop, time, stateFile).  ascii: FIXRECORDASCII
 * ("%12llu %21.13e" + " %21.13e" per field), else FIXRECORDBINARY (8-byte
 * words, exact).  state is [19][n] in TT06Dev slot order.  state_path (may
 * be NULL) receives the stateFile name. */
int cdh_checkpoint_write(const char *dir, const char *simulate_name, int loop, double time, int nx,
                         int ny, int nz, int64_t n, const int64_t *gid, const double *Vm,
                         const double *state, int ascii, char *state_path, int path_len);
/* readCheckpoint: fields by name (units converted to the model's), records
 * by gid (each of the n gids exactly once); Vm and state are in/out: a field
 * the file lacks is left as passed.  state_file: "<base>", "<base>#" or
 * "<base>#000000"; files #000001.. of a multi-file state are read in turn. */
int cdh_checkpoint_read(const char *state_file, int64_t n, const int64_t *gid, double *Vm,
                        double *state, int *loop, double *time);
/* the engine's cells through the two calls above (input cell order) */
int cdh_sim_write_checkpoint(cdh_sim *s, const char *dir, int ascii, char *state_path,
                             int path_len);
/* sets the states and Vm from the file (fields it lacks keep their
 * values); the engine's clock takes the file's loop and time (written with
 * %f, as the reference restarts from).  Sensors are not part of a
 * checkpoint: as the reference's, created anew at a restart, they start
 * fresh, their first evaluation at the step after the restart, and record
 * cells already above a level then. */
int cdh_sim_read_checkpoint(cdh_sim *s, const char *state_file);


/* ------------------------------------------------------------------------ */
/* Batched engine: many slabs advanced together, one launch per step          */
/* ------------------------------------------------------------------------ */
/* No single counterpart in the reference: each slab is what one Cardioid
 * run (simulationLoop, elec/simulationLoop.cc:150-243, on one anatomy) would
 * compute; the engine advances any number of them together.  Slabs keep
 * their own clock (time += dt from 0 when added), stimuli, drug scales and
 * sensors; their results do not depend on what else is in the batch.  Device
 * memory is one arena sized at creation (cells, box points, slabs, parameter
 * sets); slabs can be added and removed between runs. */
typedef struct cdh_batch cdh_batch;

typedef struct {
   int mode;                    /* CDH_MODE_FP64_FUSED or CDH_MODE_MIXED_FUSED */
   double dt;
   const int *fit_l;
   const int *fit_m;
   const double *fit_coef;
   int64_t cell_capacity;       /* cells; each slab takes ceil(ncell/64)*64 */
   int64_t box_capacity;        /* box points; a slab takes (bx*by*bz) of its padded box */
   int slab_capacity;
   int param_capacity;          /* parameter sets; a slab takes one per (cell type, scales) */
   int sensors;                 /* bit 0: activation times; bit 1: APD90 as well */
   int exact_functions;         /* 1: TT06Dev configuration A (tolerance 0), mode
                                   CDH_MODE_FP64_FUSED only; fit arrays unused, the
                                   slabs' gate_tail_rule ignored (no v2 loop bounds) */
   int jhTauSmooth;             /* with exact_functions: the deck's mod keyword */
} cdh_batch_params;

typedef struct {
   int nx, ny, nz;
   double dx, dy, dz, diffusionScale;
   int64_t ncell;
   const int64_t *gid;          /* ncell */
   const uint8_t *celltype;     /* ncell, index into celltypes */
   const double *sigma;         /* 6*ncell */
   int ncelltype;               /* <= 8 */
   const cdh_celltype *celltypes;
   const double *scales;        /* 3*ncell (s_kr, s_cal, s_na) or NULL */
   double threshold;            /* activation threshold, mV */
   int gate_tail_rule;
   int nstim;                   /* <= 8 */
   const cdh_box_stim *stim;
   int check_ranges;            /* as cdh_sim_params */
   double vm_min, vm_max;
   int setup_on_host;           /* as cdh_sim_params */
   /* remote cells (Diffusion's nRemote, elec/FGRDiffusionOMP.cc:14-73): tissue
    * inside the slab's padded box that another slab owns; the stencil weights
    * count them as tissue and their Vm comes from cdh_batch_link_halo */
   int64_t nremote;
   const int64_t *remote_gid;   /* nremote */
   const double *remote_sigma;  /* 6*nremote */
} cdh_slab_desc;

int cdh_batch_create(const cdh_batch_params *p, cdh_batch **out);
int cdh_batch_destroy(cdh_batch *b);
/* arena units a slab needs: cells (whole tiles of 64), box points, parameter sets */
int cdh_slab_extent(const cdh_slab_desc *d, int64_t *cells, int64_t *box_points, int *param_sets);
/* adds a slab; CDH_ERR_FULL if the arena has no room for it now */
int cdh_batch_add(cdh_batch *b, const cdh_slab_desc *d, int *slab);
int cdh_batch_remove(cdh_batch *b, int slab);
/* HaloExchange (elec/HaloExchange.hh:197-236) on one device: after every
 * step the remote cells of slab dst that slab src owns take src's new Vm (and
 * their current Vm now).  Link every (dst, src) pair of neighbours; a slab's
 * links end when either slab is removed.  *npoints: points linked. */
int cdh_batch_link_halo(cdh_batch *b, int dst, int src, int64_t *npoints);
/* checkpoint / restart of one slab, as cdh_sim_write_checkpoint /
 * cdh_sim_read_checkpoint (the slab's cells in the order they were added,
 * its own clock): a served request can be written out, removed and resumed */
int cdh_batch_write_checkpoint(cdh_batch *b, int slab, const char *dir, int ascii,
                               char *state_path, int path_len);
int cdh_batch_read_checkpoint(cdh_batch *b, int slab, const char *state_file);
/* A slab's sensors and exact clock, which a checkpoint file does not hold
 * (no counterpart in the reference: its sensors are not checkpointed and its
 * restart time is the file's %f).  rows: 6 host arrays of the slab's ncell
 * doubles, input cell order: tact0, tact, trec, Vmax, t90, APD90 phase
 * (0 waiting, 1 tracking Vmax, 2 done); rows the batch has no sensors for
 * read -1.  After a set the sensors continue the run (the step's Vm before
 * is their previous one), so write checkpoint + get sensors, remove, add +
 * read checkpoint + set sensors resumes an evicted slab exactly. */
int cdh_batch_get_sensors(cdh_batch *b, int slab, double *rows, double *time, int *loop);
int cdh_batch_set_sensors(cdh_batch *b, int slab, const double *rows, double time, int loop);
/* advances every slab by nsteps; asynchronous on the batch's stream */
int cdh_batch_run(cdh_batch *b, int nsteps);
int cdh_batch_run_timed(cdh_batch *b, int nsteps, double *ms);
int cdh_batch_sync(cdh_batch *b);
/* steps per captured HIP graph (even; 0 launches each step directly) */
int cdh_batch_set_graph(cdh_batch *b, int steps);
/* launch configuration: the kernel instance's register target (waves/SIMD:
 * FP64 2..3, mixed 3..4), the workgroup size (64, 128, 256), a cap on the
 * waves per SIMD resident at run time (0: none; applied by reserving LDS per
 * workgroup) and the order in which tiles meet the 8 XCDs, each with its own
 * L2 (-1: XCD x serves the x-th contiguous eighth of the tiles; k >= 0: runs
 * of 2^k workgroups dealt to the XCDs in turn) */
int cdh_batch_set_launch(cdh_batch *b, int waves, int block, int occupancy, int tile_order);
int cdh_batch_get_launch(cdh_batch *b, int *waves, int *block, int *occupancy, int *tile_order);
/* times every launch configuration for `steps` steps from the current state
 * (which is restored afterwards) and keeps the fastest; the table gets
 * (waves, block, occupancy, tile order, ms per step) rows, up to max_rows */
int cdh_batch_autotune(cdh_batch *b, int steps, double *best_ms, double *table, int max_rows,
                       int *nrows);
/* per-slab results, in the slab's input cell order */
int cdh_batch_get_vm(cdh_batch *b, int slab, double *vm);
int cdh_batch_get_state(cdh_batch *b, int slab, double *state);      /* [19][ncell] */
int cdh_batch_get_activation(cdh_batch *b, int slab, double *tact0, double *tact, double *trec);
int cdh_batch_get_t90(cdh_batch *b, int slab, double *t90);
/* cdh_activation_sensor_counts for one slab; reset with its sensors (add,
 * restart), set from the times by cdh_batch_set_sensors */
int cdh_batch_sensor_counts(cdh_batch *b, int slab, int64_t *n_tact0, int64_t *n_tact,
                            int64_t *n_trec);
int cdh_batch_range_report(cdh_batch *b, int slab, int64_t *count, int *loop, int64_t *cell,
                           double *V, double *dVmR, double *dVmD);
int cdh_batch_slab_info(cdh_batch *b, int slab, int64_t *ncell, double *time, int *loop);
int cdh_batch_info(cdh_batch *b, int64_t *device_bytes, int64_t *cells_used, int64_t *box_used,
                   int *nslab);

/* ---- measurement ----------------------------------------------------------
 * No counterpart in the reference: the practical ceiling for the roofline of
 * the fused FP64 step.  Times `reps` passes of one HBM stream over `cells`
 * cells of 19 FP64 fields in 4096-cell runs with the fields in pairs, one
 * non-temporal 16-byte access per lane and pair (the fused kernel's state
 * pattern): pattern 0 reads,
 * 1 writes, 2 reads and writes in place.  *gbps = bytes moved / time.
 * Allocates and frees its own buffer (cells * 152 bytes) and stream. */
int cdh_stream_rate(int pattern, int64_t cells, int reps, double *gbps);

#ifdef __cplusplus
}
#endif
#endif
