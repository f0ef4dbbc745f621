// engine_host.hip -- the multi-slab fused engine (host side); see
// engine_host.hpp.  Compiled with -ffp-contract=off so that the host copy of
// each slab's clock rounds as the device copy does.
#include "engine_host.hpp"

#include <algorithm>
#include <cstring>

#include "setup.hpp"
#include "setup_device.hpp"

namespace cdh {

// ---------------------------------------------------------------- RangeAlloc

RangeAlloc::RangeAlloc(int64_t capacity) : cap_(capacity)
{
   if (capacity > 0) free_[0] = capacity;
}

int64_t RangeAlloc::alloc(int64_t len)
{
   if (len <= 0) return -1;
   for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      const int64_t off = it->first, rest = it->second - len;
      free_.erase(it);
      if (rest > 0) free_[off + len] = rest;
      taken_[off] = len;
      used_ += len;
      return off;
   }
   return -1;
}

void RangeAlloc::release(int64_t off, int64_t len)
{
   auto t = taken_.find(off);
   if (t == taken_.end() || t->second != len) throw Error(kErrState, "RangeAlloc: bad release");
   taken_.erase(t);
   used_ -= len;
   auto next = free_.lower_bound(off);
   if (next != free_.end() && next->first == off + len) {
      len += next->second;
      free_.erase(next);
   }
   auto it = free_.lower_bound(off);
   if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == off) {
         prev->second += len;
         return;
      }
   }
   free_[off] = len;
}

int64_t RangeAlloc::end() const
{
   if (taken_.empty()) return 0;
   const auto &last = *taken_.rbegin();
   return last.first + last.second;
}

// ---------------------------------------------------------------- helpers

int mixed_slot_source(int q)
{
   static const int src[kMixed32] = {S_CA_SS, S_CA_SR, S_R_PRIME, S_FCASS, -1,
                                     S_M, S_H, S_J, S_XR1, S_XR2, S_XS, S_R, S_D, S_F, S_F2,
                                     S_JL, S_S};
   return src[q];
}

namespace {

// HaloExchange within one device: box[dst] = box[src] per (dst, src) pair;
// destinations are remote points and sources local cells, never the same
__global__ __launch_bounds__(256) void k_halo(int64_t n, const uint32_t *pairs, double *box)
{
   const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
   if (i < n) box[pairs[2 * i]] = box[pairs[2 * i + 1]];
}

using KernelFn = void (*)(EngineArgs);

KernelFn kernel_for(int mode, bool exact, int waves)
{
   if (mode == CDH_MODE_MIXED_FUSED) {
      return waves == 4 ? k_engine_mixed<4> : k_engine_mixed<3>;
   }
   if (exact) return waves == 2 ? k_engine_f64<2, kExact> : k_engine_f64<3, kExact>;
   if (mode == CDH_MODE_FP64_FMA)
      return waves == 2 ? k_engine_f64<2, kFitsFma> : k_engine_f64<3, kFitsFma>;
   return waves == 2 ? k_engine_f64<2, kFits> : k_engine_f64<3, kFits>;
}

int min_waves(int mode) { return mode == CDH_MODE_MIXED_FUSED ? 3 : 2; }
int max_waves(int mode) { return mode == CDH_MODE_MIXED_FUSED ? 4 : 3; }

// the engines' states: a slab's [nq][n] array from its device range of S
// cells in runs, one field per access (run_index) or in pairs
// (run_pair_index), and back
inline int64_t state_index(bool pairs, int q, int64_t k, int nq, int64_t S)
{
   return pairs ? run_pair_index(q, k, nq, S) : run_index(q, k, nq, S);
}
template <class T>
std::vector<T> from_runs(const std::vector<T> &b, int nq, int64_t n, int64_t S, bool pairs)
{
   std::vector<T> out((size_t)nq * n);
   for (int q = 0; q < nq; q++)
      for (int64_t k = 0; k < n; k++) out[q * n + k] = b[state_index(pairs, q, k, nq, S)];
   return out;
}

template <class T>
std::vector<T> to_runs(const std::vector<T> &soa, int nq, int64_t S, bool pairs)
{
   std::vector<T> b((size_t)nq * S);
   for (int q = 0; q < nq; q++)
      for (int64_t k = 0; k < S; k++) b[state_index(pairs, q, k, nq, S)] = soa[q * S + k];
   return b;
}

template <class T>
std::vector<T> block(const std::vector<T> &soa, int nq, int64_t S, int64_t stride)
{
   std::vector<T> b((size_t)nq * S);
   for (int q = 0; q < nq; q++)
      for (int64_t k = 0; k < S; k++) b[blk_index(q, k, nq)] = soa[q * stride + k];
   return b;
}

// the mixed engine's 9 weight words of S cells, 9 plain rows of S, into
// their load groups (grp_index9)
std::vector<uint32_t> group9(const std::vector<uint32_t> &soa, int64_t S)
{
   std::vector<uint32_t> b((size_t)9 * S);
   for (int q = 0; q < 9; q++)
      for (int64_t k = 0; k < S; k++) b[grp_index9(q, k)] = soa[q * S + k];
   return b;
}

template <class T>
void download(const T *dev, T *host, size_t count)
{
   if (count) CDH_HIP(hipMemcpy(host, dev, count * sizeof(T), hipMemcpyDeviceToHost));
}

}  // namespace

// The autotuner's axes, each of which changes what runs: the workgroup size,
// the occupancy cap, and the order in which tiles meet the XCDs (contiguous
// eighths, or runs of 1 or 16 workgroups dealt in turn).  The kernel
// instances' register targets are not an axis: with the current kernels the
// two targets of each mode compile to the same register count.
std::vector<LaunchCfg> Engine::candidates(int mode)
{
   std::vector<LaunchCfg> c;
   const int w0 = min_waves(mode);
   for (int b : {64, 128, 256})
      for (int o : {0, 3})
         for (int x : {-1, 0, 4}) c.push_back(LaunchCfg{w0, b, o, x});
   return c;
}

// LDS a workgroup reserves so that at most occ waves per SIMD (4*occ per
// CU) are resident: just under 1/(workgroups per CU) of the CU's LDS
int Engine::lds_bytes(const LaunchCfg &c) const
{
   if (c.occ <= 0) return 0;
   const int wpg = c.block / kBlk;
   const int nwg = std::max(1, 4 * c.occ / wpg);
   return std::max(0, (lds_per_cu_ / nwg - 256) / 512 * 512);
}

// ---------------------------------------------------------------- Engine

template <class T>
void Engine::track(DevBuf<T> &b, size_t count)
{
   b.alloc(count);
   bytes_ += (int64_t)(count * sizeof(T));
}

// Pinned staging for host-to-device copies on the engine's stream.  Small
// copies (the descriptors and lists of one add, up to kSmallPut bytes each)
// share a block, each taking the next 256-byte-aligned piece of it; a large
// copy takes a block of its own.  A block's event is recorded after every
// copy out of it, so once the event has completed every copy out of the
// block is done and the whole block is free again.
char *Engine::pinned(size_t bytes, hipEvent_t *ev)
{
   for (PinBlock &b : pins_)
      if (b.busy && hipEventQuery(b.ev) == hipSuccess) {
         b.busy = false;
         b.used = 0;
      }
   const size_t mib = size_t(1) << 20;
   const bool small = bytes <= kSmallPut;
   const size_t need = (bytes + 255) / 256 * 256;
   PinBlock *blk = nullptr;
   if (small && fill_ >= 0 && pins_[fill_].used + need <= pins_[fill_].cap) blk = &pins_[fill_];
   if (!blk) {
      for (PinBlock &b : pins_)
         if (!b.busy && b.cap >= need && (!blk || b.cap < blk->cap)) blk = &b;
      if (!blk) {
         const size_t cap = std::max(mib, (need + mib - 1) / mib * mib);
         if (pinned_bytes_ + cap > kPinnedLimit) return nullptr;
         PinBlock b;
         b.cap = cap;
         CDH_HIP(hipHostMalloc((void **)&b.p, cap, hipHostMallocDefault));
         if (hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipHostFree(b.p);
            throw Error(kErrHip, "engine: hipEventCreateWithFlags failed");
         }
         pins_.push_back(b);
         pinned_bytes_ += cap;
         blk = &pins_.back();
      }
      fill_ = small ? (int)(blk - pins_.data()) : (fill_ == (int)(blk - pins_.data()) ? -1 : fill_);
   }
   char *p = blk->p + blk->used;
   blk->used = small ? blk->used + need : blk->cap;
   blk->busy = true;
   *ev = blk->ev;
   return p;
}

// Keeps at most kPinnedKeep bytes of idle staging: blocks whose copies are
// done are freed, largest first, above that mark.
void Engine::trim_pinned()
{
   if (pinned_bytes_ <= kPinnedKeep) return;
   std::vector<size_t> idle;
   for (size_t k = 0; k < pins_.size(); k++) {
      PinBlock &b = pins_[k];
      if (b.busy && hipEventQuery(b.ev) == hipSuccess) {
         b.busy = false;
         b.used = 0;
      }
      if (!b.busy) idle.push_back(k);
   }
   std::sort(idle.begin(), idle.end(),
             [&](size_t a, size_t b) { return pins_[a].cap > pins_[b].cap; });
   std::vector<bool> drop(pins_.size(), false);
   for (size_t k : idle) {
      if (pinned_bytes_ <= kPinnedKeep) break;
      (void)hipHostFree(pins_[k].p);
      (void)hipEventDestroy(pins_[k].ev);
      pinned_bytes_ -= pins_[k].cap;
      drop[k] = true;
   }
   std::vector<PinBlock> kept;
   for (size_t k = 0; k < pins_.size(); k++)
      if (!drop[k]) kept.push_back(pins_[k]);
   pins_.swap(kept);
   fill_ = -1;
}

template <class T>
void Engine::put(T *dst, const std::vector<T> &src)
{
   if (src.empty()) return;
   const size_t bytes = src.size() * sizeof(T);
   hipEvent_t ev = nullptr;
   if (char *p = pinned(bytes, &ev)) {
      std::memcpy(p, src.data(), bytes);
      CDH_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, stream_));
      CDH_HIP(hipEventRecord(ev, stream_));
      return;
   }
   auto keep = std::make_shared<std::vector<T>>(src);
   CDH_HIP(hipMemcpyAsync(dst, keep->data(), keep->size() * sizeof(T), hipMemcpyHostToDevice,
                          stream_));
   staging_.push_back(keep);
   CDH_HIP(hipEventRecord(staging_ev_, stream_));
}

void Engine::refresh_tiles()
{
   if (!list_dirty_) return;
   std::vector<const Slab *> live;
   for (const Slab &s : slabs_)
      if (s.live) live.push_back(&s);
   std::sort(live.begin(), live.end(),
             [](const Slab *a, const Slab *b) { return a->tile0 < b->tile0; });
   std::vector<uint32_t> list;
   for (const Slab *s : live)
      for (int64_t t = 0; t < s->ntile; t++) list.push_back((uint32_t)(s->tile0 + t));
   nlist_ = (int64_t)list.size();
   tiles_direct_ = !list.empty() && (int64_t)(list.back() - list.front()) + 1 == nlist_;
   tile_base_ = list.empty() ? 0 : (int64_t)list.front();
   if (!tiles_direct_) put(tile_list_.p, list);
   list_dirty_ = false;
}

void Engine::retire_staging()
{
   trim_pinned();
   if (staging_.empty()) return;
   if (hipEventQuery(staging_ev_) == hipSuccess) staging_.clear();
}

Engine::Engine(int mode, double dt, const int *fit_l, const int *fit_m, const double *fit_coef,
               int64_t cap_cells, int64_t cap_box, int cap_slabs, int cap_params, int sensors,
               int exact, int jhTauSmooth)
    : mode_(mode), dt_(dt), sensors_(sensors), exact_(exact ? 1 : 0)
{
   CDH_REQUIRE(mode == CDH_MODE_FP64_FUSED || mode == CDH_MODE_FP64_FMA ||
                   mode == CDH_MODE_MIXED_FUSED,
               "engine: mode must be FP64 fused, FP64 FMA or mixed fused");
   if (exact_ && mode != CDH_MODE_FP64_FUSED)
      throw Error(kErrUnsupported,
                  "engine: the exact functions (TT06Dev tolerance 0) run in the FP64 fused mode "
                  "only");
   CDH_REQUIRE(dt > 0, "engine: dt must be positive");
   CDH_REQUIRE(exact_ || (fit_l && fit_m && fit_coef), "engine: null fit arrays");
   CDH_REQUIRE(cap_cells >= 1 && cap_cells < (int64_t(1) << 31) - 64,
               "engine: cell capacity must be 1..2^31-65");
   CDH_REQUIRE(cap_box >= 1 && cap_box < (int64_t(1) << 28), "engine: box capacity must be below 2^28");
   CDH_REQUIRE(cap_slabs >= 1 && cap_slabs < kNoSlab, "engine: slab capacity must be 1..65534");
   CDH_REQUIRE(cap_params >= 1, "engine: parameter-set capacity must be positive");
   // until a set_launch or an autotune: what the autotuner picks on the
   // 20.1 M-cell shell (one box's table: FP64 FMA 1.477 ms per step with
   // 64-thread workgroups against 1.556 with 256; mixed 0.935 with 128 and
   // runs of 16 tile groups per XCD against 0.950 with 256, contiguous)
   const bool mixed = mode == CDH_MODE_MIXED_FUSED;
   cfg_.waves = mixed ? 3 : 2;
   cfg_.block = mixed ? 128 : 64;
   cfg_.xcd = mixed ? 4 : -1;
   const int64_t ntile = (cap_cells + kBlk - 1) / kBlk;
   const int64_t S = ntile * kBlk;
   tiles_ = RangeAlloc(ntile);
   boxes_ = RangeAlloc(cap_box);
   params_ = RangeAlloc(cap_params);
   slabs_.resize(cap_slabs);

   int dev = 0;
   CDH_HIP(hipGetDevice(&dev));
   CDH_HIP(hipDeviceGetAttribute(&lds_per_cu_, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                 dev));
   CDH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
   CDH_HIP(hipEventCreate(&ev0_));
   CDH_HIP(hipEventCreate(&ev1_));
   CDH_HIP(hipEventCreateWithFlags(&staging_ev_, hipEventDisableTiming));

   tt06_constants(&K_);
   track(dK_, 1);
   dK_.upload(&K_, 1);
   std::vector<double> coef = exact_ ? std::vector<double>(kFitCoefTotal, 0.0)
                                     : scaled_fit_coefficients(fit_l, fit_m, fit_coef, dt);
   if (exact_) {
      TT06ExactConst E;
      tt06_exact_constants(&E, jhTauSmooth);
      track(dE_, 1);
      dE_.upload(&E, 1);
   }
   track(coef_, coef.size());
   coef_.upload(coef.data(), coef.size());
   track(psets_, cap_params);
   track(meta_, S);
   track(bi_, S);
   track(blk_slab_, ntile);
   track(tile_list_, ntile);
   CDH_HIP(hipMemset(blk_slab_.p, 0xFF, ntile * sizeof(uint16_t)));
   track(slabdev_, cap_slabs);
   slabdev_.zero();
   track(ranges_, cap_slabs);
   ranges_.zero();
   track(counts_, 4 * (size_t)cap_slabs);
   counts_.zero();
   if (mode != CDH_MODE_MIXED_FUSED) {
      track(state_, (size_t)N_STATE * S);
      track(Aw_, (size_t)18 * S);
   } else {
      track(st32_, (size_t)kMixed32 * S);
      track(st64_, (size_t)2 * S);
      track(Aref_, (size_t)18 * ntile);
      track(Aw16_, (size_t)9 * S);
   }
   for (int k = 0; k < 2; k++) {
      track(box_[k], cap_box);
      box_[k].zero();
   }
   if (sensors & 1) {
      track(tact0_, S);
      track(tact_, S);
      track(trec_, S);
   }
   if (sensors & 2) {
      track(apd_phase_, S);
      track(vmax_, S);
      track(t90_, S);
   }
   CDH_HIP(hipDeviceSynchronize());
}

Engine::~Engine()
{
   if (stream_) (void)hipStreamSynchronize(stream_);
   if (graph_) (void)hipGraphExecDestroy(graph_);
   if (ev0_) (void)hipEventDestroy(ev0_);
   if (ev1_) (void)hipEventDestroy(ev1_);
   if (staging_ev_) (void)hipEventDestroy(staging_ev_);
   for (PinBlock &b : pins_) {
      (void)hipHostFree(b.p);
      (void)hipEventDestroy(b.ev);
   }
   if (stream_) (void)hipStreamDestroy(stream_);
}

bool Engine::fits(int64_t ntile, int64_t nbox, int64_t nparam) const
{
   auto room = [](const RangeAlloc &a, int64_t len) {
      RangeAlloc c = a;
      return c.alloc(len) >= 0;
   };
   bool slot = false;
   for (const Slab &s : slabs_) slot |= !s.live;
   return slot && room(tiles_, ntile) && room(boxes_, nbox) && room(params_, nparam);
}

int Engine::nslab() const
{
   int k = 0;
   for (const Slab &s : slabs_) k += s.live;
   return k;
}

void Engine::check_slab(int id) const
{
   CDH_REQUIRE(id >= 0 && id < (int)slabs_.size() && slabs_[id].live, "engine: no such slab");
}

const Engine::Slab &Engine::slab(int id) const
{
   check_slab(id);
   return slabs_[id];
}

int Engine::add(const SlabHost &h)
{
   retire_staging();
   int id = -1;
   for (int k = 0; k < (int)slabs_.size(); k++)
      if (!slabs_[k].live) {
         id = k;
         break;
      }
   if (id < 0) return -1;
   const int64_t t0 = tiles_.alloc(h.ntile);
   const int64_t b0 = t0 < 0 ? -1 : boxes_.alloc(h.nbox);
   const int64_t p0 = b0 < 0 ? -1 : params_.alloc((int64_t)h.psets.size());
   if (t0 < 0 || b0 < 0 || p0 < 0) {
      if (b0 >= 0) boxes_.release(b0, h.nbox);
      if (t0 >= 0) tiles_.release(t0, h.ntile);
      return -1;
   }
   // a failure part way (a device allocation, an upload) gives the slab's
   // ranges back and leaves no half-built slab behind
   try {
      const int64_t S = h.S, c0 = t0 * kBlk;
      Slab &s = slabs_[id];
      s = Slab{};
      s.live = true;
      s.tile0 = t0;
      s.ntile = h.ntile;
      s.box0 = b0;
      s.nbox = h.nbox;
      s.param0 = p0;
      s.nparam = (int64_t)h.psets.size();
      s.n = h.n;
      s.geom = h.geom;
      s.perm = h.perm;
      s.bi.assign(h.bi.begin(), h.bi.begin() + h.n);
      s.K0.resize(h.n);
      for (int64_t k = 0; k < h.n; k++) s.K0[k] = h.psets[h.meta[k] & 0xFFu].K0;
      s.sensors = sensors_;
      s.gnx = h.gnx;
      s.gny = h.gny;
      s.remote_gid = h.remote_gid;
      s.remote_bi = h.remote_bi;

      put(psets_.p + p0, h.psets);
      put(meta_.p + c0, h.meta);
      std::vector<uint32_t> bi(S);
      for (int64_t k = 0; k < S; k++) bi[k] = (uint32_t)(b0 + h.bi[k]);
      put(bi_.p + c0, bi);
      if (mode_ != CDH_MODE_MIXED_FUSED) {
         put(state_.p + c0 * N_STATE, to_runs(h.state, N_STATE, S, true));
      } else {
         std::vector<float> s32((size_t)kMixed32 * S);
         std::vector<double> s64((size_t)2 * S);
         for (int64_t k = 0; k < S; k++) {
            for (int q = 0; q < kMixed32; q++) {
               const int src = mixed_slot_source(q);
               s32[q * S + k] = src < 0 ? 0.0f : (float)h.state[src * S + k];
            }
            s64[k] = h.state[S_CA_I * S + k];
            s64[S + k] = h.state[S_NA_I * S + k];
         }
         put(st32_.p + c0 * kMixed32, to_runs(s32, kMixed32, S, false));
         put(st64_.p + c0 * 2, to_runs(s64, 2, S, false));
      }
      if (h.device_weights) {
         // the weights are built below, after the slab's descriptor, which
         // the set-up writes the compact weights' scale into
      } else if (mode_ != CDH_MODE_MIXED_FUSED) {
         put(Aw_.p + c0 * 18, block(h.Aw, 18, S, S));
      } else {
         put(Aref_.p + t0 * 18, h.Aref);
         put(Aw16_.p + c0 * 9, group9(h.Aw16, S));
      }
      put(box_[0].p + b0, h.box0);
      put(box_[1].p + b0, h.box0);
      reset_sensors(s);
      put(counts_.p + 4 * id, std::vector<uint32_t>(4, 0u));
      SlabDev d{};
      d.time[parity_] = 0.0;
      d.loop[parity_] = 0;
      d.active = 1;
      d.first_blk = (int32_t)t0;
      d.ntile = (int32_t)h.ntile;
      for (int k = 0; k < 19; k++) d.soff[k] = (uint32_t)((h.off[k] + kBoxShift) * 8);
      d.param_base = (int32_t)p0;
      d.nstim = (int32_t)h.stim.size();
      d.scale = h.diffusionScale;
      d.threshold = h.threshold;
      d.check_range = h.check_ranges;
      d.wdelta_scale = h.wdelta_scale;
      d.sensor_start = 0;
      d.vm_min = h.vm_min;
      d.vm_max = h.vm_max;
      put(ranges_.p + id, std::vector<RangeRec>(1, RangeRec{}));
      for (int q = 0; q < d.nstim; q++) {
         const cdh_box_stim &b = h.stim[q];
         d.stim[q] = PulseDev{b.t0, b.tf, b.period, b.phase, b.duration, b.amplitude};
      }
      slab_stim_eval(d.stim, d.nstim, 0.0, &d.stim_on[parity_], d.stim_val[parity_]);
      put(slabdev_.p + id, std::vector<SlabDev>{d});
      if (h.device_weights) {
         // everything in stream order: no wait for the device
         const bool f64 = mode_ != CDH_MODE_MIXED_FUSED;
         const int64_t nremote = (int64_t)h.remote_bi.size();
         setup_.reserve(h.nbox, h.n, nremote, S, !f64);
         put(setup_.sigc.p, h.sig_cells);
         if (nremote > 0) {
            put(setup_.rbi.p, h.remote_bi);
            put(setup_.rsig.p, h.remote_sig);
         }
         device_slab_weights(stream_, setup_, h.geom, h.nbox, h.n, h.ntile, bi_.p + c0,
                             (uint32_t)b0, nremote, f64 ? Aw_.p + c0 * 18 : nullptr,
                             f64 ? nullptr : Aref_.p + t0 * 18, f64 ? nullptr : Aw16_.p + c0 * 9,
                             &slabdev_.p[id].wdelta_scale);
      }
      put(blk_slab_.p + t0, std::vector<uint16_t>(h.ntile, (uint16_t)id));
   } catch (...) {
      slabs_[id] = Slab{};
      params_.release(p0, (int64_t)h.psets.size());
      boxes_.release(b0, h.nbox);
      tiles_.release(t0, h.ntile);
      throw;
   }
   list_dirty_ = true;
   return id;
}

void Engine::remove(int id)
{
   check_slab(id);
   retire_staging();
   list_dirty_ = true;
   const size_t nl = links_.size();
   links_.erase(std::remove_if(links_.begin(), links_.end(),
                               [&](const HaloLink &l) { return l.dst == id || l.src == id; }),
                links_.end());
   if (links_.size() != nl) upload_halo();
   Slab &s = slabs_[id];
   put(blk_slab_.p + s.tile0, std::vector<uint16_t>(s.ntile, kNoSlab));
   put(&slabdev_.p[id].active, std::vector<int32_t>{0});
   tiles_.release(s.tile0, s.ntile);
   boxes_.release(s.box0, s.nbox);
   params_.release(s.param0, s.nparam);
   s = Slab{};
}

EngineArgs Engine::args(int parity) const
{
   EngineArgs a{};
   a.K = dK_.p;
   a.coef = coef_.p;
   a.params = psets_.p;
   a.meta = meta_.p;
   a.bi = bi_.p;
   a.blk_slab = blk_slab_.p;
   a.slabs = slabdev_.p;
   a.state = state_.p;
   a.st32 = st32_.p;
   a.st64 = st64_.p;
   a.Aw = Aw_.p;
   a.Aref = Aref_.p;
   a.Aw16 = Aw16_.p;
   a.box_cur = box_[parity].p;
   a.box_next = box_[parity ^ 1].p;
   a.parity = parity;
   a.tiles = tiles_direct_ ? nullptr : tile_list_.p;
   a.tile_base = (int)tile_base_;
   a.ntile = (int)nlist_;
   const int wpg = cfg_.block / kBlk;
   const int ngroup = (a.ntile + wpg - 1) / wpg;
   a.xcd_chunk_log2 = cfg_.xcd;
   if (cfg_.xcd < 0) {
      a.groups_per_xcd = (ngroup + 7) / 8;
   } else {
      // whole rounds of 8 runs of G groups: every group index below
      // 8 * G * rounds is reached, the ones past ngroup exit
      const int G = 1 << cfg_.xcd;
      a.groups_per_xcd = (ngroup + 8 * G - 1) / (8 * G) * G;
   }
   a.dt = dt_;
   a.gate_mask = 0x1FFFu;
   a.sensors = sensors_;
   a.tact0 = tact0_.p;
   a.tact = tact_.p;
   a.trec = trec_.p;
   a.apd_phase = apd_phase_.p;
   a.vmax = vmax_.p;
   a.t90 = t90_.p;
   a.counts = counts_.p;
   a.ranges = ranges_.p;
   a.E = dE_.p;
   return a;
}

void Engine::launch_step(int parity)
{
   const EngineArgs a = args(parity);
   if (a.ntile == 0) return;
   const dim3 grid(8 * a.groups_per_xcd), blk(cfg_.block);
   hipLaunchKernelGGL(kernel_for(mode_, exact_ != 0, cfg_.waves), grid, blk, lds_bytes(cfg_),
                      stream_, a);
   CDH_LAUNCH_CHECK();
   if (nhalo_ > 0) {
      hipLaunchKernelGGL(k_halo, dim3((unsigned)((nhalo_ + 255) / 256)), dim3(256), 0, stream_,
                         nhalo_, halo_.p, box_[parity ^ 1].p);
      CDH_LAUNCH_CHECK();
   }
}

void Engine::upload_halo()
{
   std::vector<uint32_t> all;
   for (const HaloLink &l : links_) all.insert(all.end(), l.pairs.begin(), l.pairs.end());
   sync();
   nhalo_ = (int64_t)all.size() / 2;
   if (halo_.n < all.size()) halo_.alloc(all.size());
   put(halo_.p, all);   // in stream order, before any later step reads it
   ++halo_version_;
}

int64_t Engine::link_halo(int dst, int src)
{
   check_slab(dst);
   check_slab(src);
   CDH_REQUIRE(dst != src, "engine: a slab cannot be its own halo source");
   for (const HaloLink &l : links_)
      CDH_REQUIRE(!(l.dst == dst && l.src == src), "engine: halo already linked");
   const Slab &d = slabs_[dst], &s = slabs_[src];
   CDH_REQUIRE(d.gnx == s.gnx && d.gny == s.gny, "engine: halo slabs on different grids");
   std::vector<uint8_t> own(s.nbox, 0);
   for (uint32_t b : s.bi) own[b] = 1;
   HaloLink l{dst, src, {}};
   for (size_t r = 0; r < d.remote_gid.size(); r++) {
      const int64_t bs = box_index(s.geom, s.gnx, s.gny, d.remote_gid[r]);
      if (bs < 0 || !own[bs]) continue;
      l.pairs.push_back((uint32_t)(d.box0 + d.remote_bi[r]));
      l.pairs.push_back((uint32_t)(s.box0 + bs));
   }
   CDH_REQUIRE(!l.pairs.empty(), "engine: slab src owns none of slab dst's remote cells");
   const int64_t np = (int64_t)l.pairs.size() / 2;
   // the remote points take src's current Vm now, then after every step
   DevBuf<uint32_t> once(l.pairs.size());
   put(once.p, l.pairs);   // in stream order, before the kernel that reads it
   hipLaunchKernelGGL(k_halo, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, stream_, np,
                      once.p, box_[parity_].p);
   CDH_LAUNCH_CHECK();
   sync();                 // `once` is freed on return
   links_.push_back(std::move(l));
   upload_halo();
   return np;
}

void Engine::step_eager()
{
   launch_step(parity_);
   parity_ ^= 1;
   for (Slab &s : slabs_)
      if (s.live) {
         s.time = s.time + dt_;
         ++s.loop;
      }
}

void Engine::drop_graph()
{
   if (graph_) (void)hipGraphExecDestroy(graph_);
   graph_ = nullptr;
   graph_ntile_ = -1;
}

void Engine::rebuild_graph()
{
   drop_graph();
   hipGraph_t g = nullptr;
   CDH_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
   try {
      for (int k = 0; k < graph_steps_; k++) launch_step(k & 1);
   } catch (...) {
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
   }
   CDH_HIP(hipStreamEndCapture(stream_, &g));
   hipError_t e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
   (void)hipGraphDestroy(g);
   CDH_HIP(e);
   graph_ntile_ = (int)nlist_;
   graph_direct_ = tiles_direct_;
   graph_base_ = tile_base_;
   graph_cfg_ = cfg_;
   graph_halo_ = halo_version_;
}

void Engine::set_graph_steps(int k)
{
   CDH_REQUIRE(k >= 0 && k % 2 == 0, "engine: graph steps must be even (0 disables graphs)");
   graph_steps_ = k;
   drop_graph();
}

void Engine::set_launch(const LaunchCfg &c)
{
   CDH_REQUIRE(c.waves >= min_waves(mode_) && c.waves <= max_waves(mode_),
               "engine: no kernel instance for that register target");
   CDH_REQUIRE(c.block == 64 || c.block == 128 || c.block == 256,
               "engine: workgroup size must be 64, 128 or 256");
   CDH_REQUIRE(c.occ >= 0 && c.occ <= 8, "engine: occupancy cap must be 0..8");
   CDH_REQUIRE(c.xcd >= -1 && c.xcd <= 10, "engine: tile order must be -1 or 0..10");
   cfg_ = c;
   drop_graph();
}

void Engine::run(int nsteps)
{
   CDH_REQUIRE(nsteps >= 0, "engine: negative step count");
   retire_staging();
   refresh_tiles();      // before any capture: an upload must not land in a graph
   while (nsteps > 0) {
      if (graph_steps_ > 0 && parity_ == 0 && nsteps >= graph_steps_) {
         if (!graph_ || graph_ntile_ != (int)nlist_ || graph_direct_ != tiles_direct_ ||
             graph_base_ != tile_base_ || graph_cfg_.waves != cfg_.waves ||
             graph_cfg_.block != cfg_.block || graph_cfg_.occ != cfg_.occ ||
             graph_cfg_.xcd != cfg_.xcd ||
             graph_halo_ != halo_version_)
            rebuild_graph();
         CDH_HIP(hipGraphLaunch(graph_, stream_));
         for (Slab &s : slabs_)
            if (s.live)
               for (int k = 0; k < graph_steps_; k++) {
                  s.time = s.time + dt_;
                  ++s.loop;
               }
         nsteps -= graph_steps_;
      } else {
         step_eager();
         --nsteps;
      }
   }
}

double Engine::run_timed(int nsteps)
{
   CDH_HIP(hipEventRecord(ev0_, stream_));
   run(nsteps);
   CDH_HIP(hipEventRecord(ev1_, stream_));
   CDH_HIP(hipEventSynchronize(ev1_));
   float ms = 0.0f;
   CDH_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
   return ms;
}

void Engine::sync()
{
   CDH_HIP(hipStreamSynchronize(stream_));
   retire_staging();
}

double Engine::autotune(int steps, std::vector<std::pair<LaunchCfg, double>> *table)
{
   CDH_REQUIRE(steps >= 1, "engine: autotune needs at least one step per candidate");
   sync();
   // snapshot everything a step writes, restore it after each candidate
   const int64_t S = tiles_.end() * kBlk, nb = boxes_.end();
   std::vector<std::pair<void *, size_t>> regions;
   auto add = [&](void *p, size_t bytes) {
      if (p && bytes) regions.emplace_back(p, bytes);
   };
   add(state_.p, sizeof(double) * N_STATE * S);
   add(st32_.p, sizeof(float) * kMixed32 * S);
   add(st64_.p, sizeof(double) * 2 * S);
   add(box_[0].p, sizeof(double) * nb);
   add(box_[1].p, sizeof(double) * nb);
   add(tact0_.p, sizeof(double) * S);
   add(tact_.p, sizeof(double) * S);
   add(trec_.p, sizeof(double) * S);
   add(apd_phase_.p, S);
   add(vmax_.p, sizeof(double) * S);
   add(t90_.p, sizeof(double) * S);
   add(slabdev_.p, sizeof(SlabDev) * slabs_.size());
   add(ranges_.p, sizeof(RangeRec) * slabs_.size());
   add(counts_.p, sizeof(uint32_t) * 4 * slabs_.size());
   std::vector<std::unique_ptr<DevBuf<uint8_t>>> saved;
   for (auto &r : regions) {
      saved.emplace_back(new DevBuf<uint8_t>(r.second));
      CDH_HIP(hipMemcpyAsync(saved.back()->p, r.first, r.second, hipMemcpyDeviceToDevice, stream_));
   }
   const int par0 = parity_;
   std::vector<Slab> host0 = slabs_;
   const LaunchCfg keep = cfg_;
   const int gsteps = graph_steps_;
   graph_steps_ = 0;
   // two interleaved rounds, each candidate's faster time kept: the
   // candidates differ by a few per cent, about the spread of one short timing
   const std::vector<LaunchCfg> cand = candidates(mode_);
   std::vector<double> ms(cand.size(), -1.0);
   for (int round = 0; round < 2; round++)
      for (size_t c = 0; c < cand.size(); c++) {
         set_launch(cand[c]);
         run(1);
         const double t = run_timed(steps) / steps;
         if (ms[c] < 0 || t < ms[c]) ms[c] = t;
         for (size_t k = 0; k < regions.size(); k++)
            CDH_HIP(hipMemcpyAsync(regions[k].first, saved[k]->p, regions[k].second,
                                   hipMemcpyDeviceToDevice, stream_));
         parity_ = par0;
         slabs_ = host0;
      }
   double best = -1.0;
   LaunchCfg bestc = keep;
   for (size_t c = 0; c < cand.size(); c++) {
      if (table) table->emplace_back(cand[c], ms[c]);
      if (best < 0 || ms[c] < best) {
         best = ms[c];
         bestc = cand[c];
      }
   }
   CDH_HIP(hipStreamSynchronize(stream_));
   set_launch(bestc);
   graph_steps_ = gsteps;
   return best;
}

// ---------------------------------------------------------------- accessors

std::vector<double> Engine::vm(int id)
{
   check_slab(id);
   sync();
   const Slab &s = slabs_[id];
   std::vector<double> box(s.nbox);
   download(box_[parity_].p + s.box0, box.data(), s.nbox);
   std::vector<double> v(s.n);
   for (int64_t k = 0; k < s.n; k++) v[k] = box[s.bi[k]];
   return v;
}

void Engine::set_vm(int id, const std::vector<double> &v)
{
   check_slab(id);
   sync();
   const Slab &s = slabs_[id];
   std::vector<double> box(s.nbox);
   download(box_[parity_].p + s.box0, box.data(), s.nbox);
   for (int64_t k = 0; k < s.n; k++) box[s.bi[k]] = v[k];
   put(box_[parity_].p + s.box0, box);
   sync();
}

std::vector<double> Engine::state19(int id)
{
   check_slab(id);
   sync();
   const Slab &s = slabs_[id];
   const int64_t S = s.ntile * kBlk, c0 = s.tile0 * kBlk, n = s.n;
   if (mode_ != CDH_MODE_MIXED_FUSED) {
      std::vector<double> b((size_t)N_STATE * S);
      download(state_.p + c0 * N_STATE, b.data(), b.size());
      return from_runs(b, N_STATE, n, S, true);
   }
   std::vector<float> b32((size_t)kMixed32 * S);
   std::vector<double> b64((size_t)2 * S);
   download(st32_.p + c0 * kMixed32, b32.data(), b32.size());
   download(st64_.p + c0 * 2, b64.data(), b64.size());
   const std::vector<float> s32 = from_runs(b32, kMixed32, n, S, false);
   const std::vector<double> s64 = from_runs(b64, 2, n, S, false);
   const std::vector<double> v = vm(id);
   std::vector<double> st((size_t)N_STATE * n, 0.0);
   for (int64_t k = 0; k < n; k++) {
      for (int q = 0; q < kMixed32; q++) {
         const int d = mixed_slot_source(q);
         if (d >= 0) st[d * n + k] = s32[q * n + k];
      }
      st[S_CA_I * n + k] = s64[k];
      st[S_NA_I * n + k] = s64[n + k];
      const double K_i = s.K0[k] + (double)s32[4 * n + k];
      st[S_DVK_I * n + k] = K_i / K_.c9 + v[k];
   }
   return st;
}

void Engine::set_state19(int id, const std::vector<double> &st, const std::vector<double> &v)
{
   check_slab(id);
   sync();
   const Slab &s = slabs_[id];
   const int64_t S = s.ntile * kBlk, c0 = s.tile0 * kBlk, n = s.n;
   // padding cells keep their states: start from the device copy
   if (mode_ != CDH_MODE_MIXED_FUSED) {
      std::vector<double> b((size_t)N_STATE * S);
      download(state_.p + c0 * N_STATE, b.data(), b.size());
      for (int q = 0; q < N_STATE; q++)
         for (int64_t k = 0; k < n; k++) b[run_pair_index(q, k, N_STATE, S)] = st[q * n + k];
      put(state_.p + c0 * N_STATE, b);
   } else {
      std::vector<float> b32((size_t)kMixed32 * S);
      std::vector<double> b64((size_t)2 * S);
      download(st32_.p + c0 * kMixed32, b32.data(), b32.size());
      download(st64_.p + c0 * 2, b64.data(), b64.size());
      for (int64_t k = 0; k < n; k++) {
         for (int q = 0; q < kMixed32; q++) {
            const int d = mixed_slot_source(q);
            if (d >= 0) b32[run_index(q, k, kMixed32, S)] = (float)st[d * n + k];
         }
         const double K_i = K_.c9 * (st[S_DVK_I * n + k] - v[k]);
         b32[run_index(4, k, kMixed32, S)] = (float)(K_i - s.K0[k]);
         b64[run_index(0, k, 2, S)] = st[S_CA_I * n + k];
         b64[run_index(1, k, 2, S)] = st[S_NA_I * n + k];
      }
      put(st32_.p + c0 * kMixed32, b32);
      put(st64_.p + c0 * 2, b64);
   }
   sync();
}

void Engine::activation(int id, double *tact0, double *tact, double *trec)
{
   check_slab(id);
   CDH_REQUIRE(sensors_ & 1, "engine: created without activation sensors");
   sync();
   const Slab &s = slabs_[id];
   const int64_t c0 = s.tile0 * kBlk;
   if (tact0) download(tact0_.p + c0, tact0, s.n);
   if (tact) download(tact_.p + c0, tact, s.n);
   if (trec) download(trec_.p + c0, trec, s.n);
}

void Engine::sensor_counts(int id, int64_t *n_tact0, int64_t *n_tact, int64_t *n_trec)
{
   check_slab(id);
   CDH_REQUIRE(sensors_ & 1, "engine: created without activation sensors");
   sync();
   uint32_t c[4];
   download(counts_.p + 4 * id, c, 4);
   if (n_tact0) *n_tact0 = c[0];
   if (n_tact) *n_tact = c[1];
   if (n_trec) *n_trec = c[2];
}

// fresh sensors over the slab's tiles: no activation, no APD90 yet
void Engine::reset_sensors(const Slab &s)
{
   const int64_t c0 = s.tile0 * kBlk, S = s.ntile * kBlk;
   if (sensors_ & 1) {
      const std::vector<double> neg(S, -1.0);
      put(tact0_.p + c0, neg);
      put(tact_.p + c0, neg);
      put(trec_.p + c0, neg);
   }
   if (sensors_ & 2) {
      put(apd_phase_.p + c0, std::vector<uint8_t>(S, 0));
      put(vmax_.p + c0, std::vector<double>(S, 0.0));
      put(t90_.p + c0, std::vector<double>(S, -1.0));
   }
}

void Engine::write_clock(int id, double time, int loop, int sensor_start)
{
   CDH_REQUIRE(loop >= 0, "engine: negative loop");
   sync();
   SlabDev d{};
   download(slabdev_.p + id, &d, 1);
   d.time[parity_] = time;
   d.loop[parity_] = loop;
   d.sensor_start = sensor_start;
   slab_stim_eval(d.stim, d.nstim, time, &d.stim_on[parity_], d.stim_val[parity_]);
   put(slabdev_.p + id, std::vector<SlabDev>{d});
   sync();
   slabs_[id].time = time;
   slabs_[id].loop = loop;
}

void Engine::set_clock(int id, double time, int loop)
{
   check_slab(id);
   sync();
   reset_sensors(slabs_[id]);
   put(counts_.p + 4 * id, std::vector<uint32_t>(4, 0u));
   write_clock(id, time, loop, loop);
}

void Engine::sensor_state(int id, double *rows, double *time, int *loop)
{
   check_slab(id);
   sync();
   const Slab &s = slabs_[id];
   const int64_t c0 = s.tile0 * kBlk, n = s.n;
   for (int64_t k = 0; k < kSensorFields * n; k++) rows[k] = -1.0;
   if (sensors_ & 1) {
      download(tact0_.p + c0, rows, n);
      download(tact_.p + c0, rows + n, n);
      download(trec_.p + c0, rows + 2 * n, n);
   }
   if (sensors_ & 2) {
      download(vmax_.p + c0, rows + 3 * n, n);
      download(t90_.p + c0, rows + 4 * n, n);
      std::vector<uint8_t> ph(n);
      download(apd_phase_.p + c0, ph.data(), n);
      for (int64_t k = 0; k < n; k++) rows[5 * n + k] = ph[k];
   }
   SlabDev d{};
   download(slabdev_.p + id, &d, 1);
   *time = d.time[parity_];
   *loop = d.loop[parity_];
}

void Engine::set_sensor_state(int id, const double *rows, double time, int loop)
{
   check_slab(id);
   sync();
   const Slab &s = slabs_[id];
   const int64_t c0 = s.tile0 * kBlk, n = s.n;
   if (sensors_ & 1) {
      put(tact0_.p + c0, std::vector<double>(rows, rows + n));
      put(tact_.p + c0, std::vector<double>(rows + n, rows + 2 * n));
      put(trec_.p + c0, std::vector<double>(rows + 2 * n, rows + 3 * n));
      // the counts the step kernels keep, for the times as set
      std::vector<uint32_t> c(4, 0u);
      for (int q = 0; q < 3; q++)
         for (int64_t k = 0; k < n; k++) c[q] += rows[q * n + k] >= 0.0;
      put(counts_.p + 4 * id, c);
   }
   if (sensors_ & 2) {
      std::vector<uint8_t> ph(n);
      for (int64_t k = 0; k < n; k++) {
         const double v = rows[5 * n + k];
         CDH_REQUIRE(v == 0.0 || v == 1.0 || v == 2.0, "engine: APD phase must be 0, 1 or 2");
         ph[k] = (uint8_t)v;
      }
      put(vmax_.p + c0, std::vector<double>(rows + 3 * n, rows + 4 * n));
      put(t90_.p + c0, std::vector<double>(rows + 4 * n, rows + 5 * n));
      put(apd_phase_.p + c0, ph);
   }
   write_clock(id, time, loop, -1);
}

RangeRec Engine::range_report(int id)
{
   check_slab(id);
   sync();
   RangeRec r{};
   download(ranges_.p + id, &r, 1);
   put(ranges_.p + id, std::vector<RangeRec>(1, RangeRec{}));
   sync();
   if (r.taken) r.cell -= slabs_[id].tile0 * kBlk;
   return r;
}

void Engine::t90(int id, double *out)
{
   check_slab(id);
   CDH_REQUIRE(sensors_ & 2, "engine: created without APD90");
   sync();
   const Slab &s = slabs_[id];
   download(t90_.p + s.tile0 * kBlk, out, s.n);
}

}  // namespace cdh
