// engine_host.hpp -- the multi-slab fused engine (host side).
//
// Owns one arena of device memory sized at creation: 64-cell tiles, box
// points, slab records and parameter sets.  Slabs are added and removed
// while others run (first-fit ranges; every transfer is ordered on the
// engine's stream), all active slabs advance together, one launch per step,
// and a run of steps can be replayed from a captured HIP graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <vector>

#include "engine.hpp"
#include "hip_util.hpp"
#include "setup_device.hpp"
#include "slab.hpp"

namespace cdh {

// first-fit allocator of [0, capacity) ranges
class RangeAlloc {
 public:
   explicit RangeAlloc(int64_t capacity = 0);
   int64_t alloc(int64_t len);            // -1 if no range fits
   void release(int64_t off, int64_t len);
   int64_t end() const;                   // one past the highest allocated unit
   int64_t used() const { return used_; }
   int64_t capacity() const { return cap_; }

 private:
   std::map<int64_t, int64_t> free_;      // offset -> length
   std::map<int64_t, int64_t> taken_;
   int64_t cap_ = 0, used_ = 0;
};

struct LaunchCfg {
   int waves = 0;     // register target of the kernel instance (waves per SIMD)
   int block = 256;   // threads per workgroup: 64, 128 or 256
   int occ = 0;       // waves per SIMD allowed at run time (0: no cap), set
                      // by reserving LDS per workgroup
   int xcd = -1;      // tile order over the XCDs: -1 contiguous eighths, k >= 0
                      // runs of 2^k workgroups dealt in turn (engine_tile_group)
};

class Engine {
 public:
   struct Slab {
      bool live = false;
      int64_t tile0 = 0, ntile = 0, box0 = 0, nbox = 0, param0 = 0, nparam = 0;
      int64_t n = 0;
      BoxGeom geom{};
      std::vector<int64_t> perm;
      std::vector<uint32_t> bi;        // n: box index within the slab's box
      std::vector<double> K0;          // n: rest K_i (mixed state conversion)
      int gnx = 0, gny = 0;            // the anatomy grid's x, y sizes (gid layout)
      // remote cells: tissue box points the slab reads but does not own
      std::vector<int64_t> remote_gid;
      std::vector<uint32_t> remote_bi; // box index within the slab's box
      int sensors = 0;
      double time = 0.0;               // host copy of the slab clock
      int loop = 0;
   };

   // mode: CDH_MODE_FP64_FUSED, CDH_MODE_FP64_FMA or CDH_MODE_MIXED_FUSED.
   // exact: configuration A (TT06Dev tolerance 0: the exact functions; mod
   // jhTauSmooth), FP64 fused only; the fit arrays are then not used and
   // slabs must carry no v2 loop-bound gate flags
   Engine(int mode, double dt, const int *fit_l, const int *fit_m, const double *fit_coef,
          int64_t cap_cells, int64_t cap_box, int cap_slabs, int cap_params, int sensors,
          int exact = 0, int jhTauSmooth = 0);
   ~Engine();
   Engine(const Engine &) = delete;
   Engine &operator=(const Engine &) = delete;

   // returns the slab id, or -1 if the arena has no room for it
   int add(const SlabHost &h);
   void remove(int id);
   // Halo exchange: after every step, the remote cells of slab dst that slab
   // src owns take src's new Vm (the one-GPU form of HaloExchange); also
   // copies their current Vm now.  Returns the number of points linked.
   int64_t link_halo(int dst, int src);
   int64_t halo_points() const { return nhalo_; }
   bool fits(int64_t ntile, int64_t nbox, int64_t nparam) const;

   void run(int nsteps);
   double run_timed(int nsteps);          // ms between events around the steps
   int lds_bytes(const LaunchCfg &c) const;
   void sync();
   void set_launch(const LaunchCfg &c);
   LaunchCfg launch() const { return cfg_; }
   void set_graph_steps(int k);           // 0: launch each step; else replay k-step graphs
   // times each launch configuration over `steps` steps (after a warm-up
   // step) from the same starting state, keeps the fastest; returns its ms
   double autotune(int steps, std::vector<std::pair<LaunchCfg, double>> *table);

   const Slab &slab(int id) const;
   std::vector<double> vm(int id);                        // slab order
   std::vector<double> state19(int id);                   // [19][n], slab order
   void set_state19(int id, const std::vector<double> &st, const std::vector<double> &vm);
   void set_vm(int id, const std::vector<double> &vm);
   // restart (readCheckpoint and the SIMULATE object's loop and time): the
   // slab's clock and stimulus values at time/loop, and fresh sensors, as the
   // reference creates them anew, whose first evaluation is this step's
   void set_clock(int id, double time, int loop);
   // the sensor arrays of a slab, slab order: kSensorFields rows of n
   // doubles (tact0, tact, trec, vmax, t90, APD phase) with the clock
   void sensor_state(int id, double *rows, double *time, int *loop);
   // sets them and the clock exactly; the sensors then continue the run
   // (the step's Vm before is their previous one)
   void set_sensor_state(int id, const double *rows, double time, int loop);
   void activation(int id, double *tact0, double *tact, double *trec);   // slab order
   // cells of the slab whose tact0, tact and trec have been recorded, kept
   // on the device by the step kernels (one wavefront ballot and atomic per
   // kind that occurs); the sensors' own counts, reset with them
   void sensor_counts(int id, int64_t *n_tact0, int64_t *n_tact, int64_t *n_trec);
   void t90(int id, double *out);
   // out-of-range record since the last call (cell as a slab-order index);
   // clears it
   RangeRec range_report(int id);
   static constexpr int kSensorFields = 6;

   int mode() const { return mode_; }
   bool exact() const { return exact_ != 0; }
   int64_t device_bytes() const { return bytes_; }
   int64_t cells_used() const { return tiles_.used() * kBlk; }
   int64_t box_used() const { return boxes_.used(); }
   int nslab() const;
   hipStream_t stream() const { return stream_; }
   static std::vector<LaunchCfg> candidates(int mode);

 private:
   void step_eager();
   void launch_step(int parity);
   void upload_halo();
   void rebuild_graph();
   void drop_graph();
   EngineArgs args(int parity) const;
   void retire_staging();
   void refresh_tiles();
   // async H2D: through a pinned staging block (the host returns at once,
   // also while the stream is busy), else through a kept pageable copy
   template <class T>
   void put(T *dst, const std::vector<T> &src);
   char *pinned(size_t bytes, hipEvent_t *ev);
   void trim_pinned();
   template <class T>
   void track(DevBuf<T> &b, size_t count);
   void check_slab(int id) const;
   std::vector<double> st_download_blocked(double *dev, int nq, const Slab &s);
   void reset_sensors(const Slab &s);
   void write_clock(int id, double time, int loop, int sensor_start);

   int mode_;
   double dt_;
   int sensors_;
   hipStream_t stream_ = nullptr;
   hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
   TT06Const K_{};
   LaunchCfg cfg_;
   int parity_ = 0;
   int graph_steps_ = 0;
   hipGraphExec_t graph_ = nullptr;
   int graph_ntile_ = -1;
   LaunchCfg graph_cfg_;
   int graph_halo_ = -1;
   struct HaloLink {
      int dst, src;
      std::vector<uint32_t> pairs;     // (destination, source) arena box indices
   };
   std::vector<HaloLink> links_;
   DevBuf<uint32_t> halo_;
   int64_t nhalo_ = 0;
   int halo_version_ = 0;
   int64_t bytes_ = 0;
   int lds_per_cu_ = 0;

   RangeAlloc tiles_, boxes_, params_;
   std::vector<Slab> slabs_;
   std::vector<std::shared_ptr<void>> staging_;
   hipEvent_t staging_ev_ = nullptr;
   // pinned staging blocks, reused once the copy recorded on `ev` is done;
   // at most kPinnedLimit bytes in all
   struct PinBlock {
      char *p = nullptr;
      size_t cap = 0;
      size_t used = 0;             // bytes handed out since the block was last idle
      hipEvent_t ev = nullptr;     // recorded after every copy out of the block
      bool busy = false;
   };
   static constexpr size_t kPinnedLimit = size_t(1) << 30;
   static constexpr size_t kPinnedKeep = size_t(64) << 20;   // idle staging kept
   static constexpr size_t kSmallPut = size_t(64) << 10;     // copies that share a block
   std::vector<PinBlock> pins_;
   int fill_ = -1;                 // the block small copies are taking pieces of
   size_t pinned_bytes_ = 0;
   SlabSetupScratch setup_;        // device set-up buffers, kept across slabs
   // the live slabs' tiles in arena order, the step kernel's grid: rebuilt
   // (and uploaded in stream order) by run() after slabs come or go
   DevBuf<uint32_t> tile_list_;
   int64_t nlist_ = 0;
   bool list_dirty_ = true;
   bool tiles_direct_ = false;     // the live tiles are tile_base_ .. + nlist_ - 1
   int64_t tile_base_ = 0;
   bool graph_direct_ = false;     // the captured graph's tile mapping
   int64_t graph_base_ = -1;

   DevBuf<TT06Const> dK_;
   DevBuf<double> coef_;
   DevBuf<ParamSet> psets_;
   DevBuf<uint32_t> meta_, bi_;
   DevBuf<uint16_t> blk_slab_;
   DevBuf<SlabDev> slabdev_;
   DevBuf<double> state_, st64_;
   DevBuf<float> st32_, Aw_, Aref_;
   DevBuf<uint32_t> Aw16_;
   DevBuf<double> box_[2];
   DevBuf<double> tact0_, tact_, trec_, vmax_, t90_;
   DevBuf<uint8_t> apd_phase_;
   DevBuf<RangeRec> ranges_;
   DevBuf<uint32_t> counts_;      // per slab: recorded tact0, tact, trec, (unused)
   int exact_ = 0;                 // configuration A
   DevBuf<TT06ExactConst> dE_;     // its constants
};

// mixed-mode FP32 slots (tt06_mixed.hpp order) and the reference slot each
// holds; -1 is K_i - K0
int mixed_slot_source(int q);
constexpr int kMixed32 = 17;

}  // namespace cdh
