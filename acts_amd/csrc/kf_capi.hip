// kf_capi.hip -- the C ABI of include/acts_kf.h: context set-up, device
// memory, kernel launches.  Replaces the construction and the fit() call of
// Acts::KalmanFitter<Propagator<EigenStepper<>, DirectNavigator>, traj>
// (KalmanFitter.hpp:265-270, 721-731) for a whole batch of tracks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kf_field_host.hpp"
#include "kf_reference.hpp"
#include "kf_reverse.hpp"
#include "kf_states.hpp"

using namespace kf;

namespace {

thread_local std::string g_lastError;

int fail(int code, const std::string& msg) {
  g_lastError = msg;
  return code;
}

#define KF_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return fail(ACTS_KF_ERR_DEVICE,                                         \
                  std::string(#call) + ": " + hipGetErrorString(e_));         \
  } while (0)

constexpr int kBlock = 64;
// geometry kept in LDS by the forward kernel (224 B per surface)
constexpr int kLdsSurfaces = 96;
constexpr int kLdsSlabs = 32;
// the track queue counts in int32 and may overshoot n by one per lane
constexpr int64_t kMaxBatch = (int64_t{1} << 31) - (int64_t{1} << 24);

// Workspace of acts_kf_fit_batch_device_noalloc: a fixed block of device
// memory that the library owns on every device it runs on, for callers that
// cannot allocate.  Every call on a device shares that device's copy (calls
// that overlap on one device race).  Host code never passes the array's name
// to a kernel: it uses the device address hipGetSymbolAddress returns for
// the current device, fetched once per device (staticWorkspace).
constexpr size_t kStaticWorkspaceBytes = size_t{256} << 20;
static __device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];
constexpr int kMaxDevices = 64;

#ifndef KF_FORWARD_WAVES_PER_EU
#define KF_FORWARD_WAVES_PER_EU 1
#endif

// Forward filter, persistent: each wavefront keeps its 64 lanes busy with
// tracks taken from a global queue (`next`), in the order `order` when given
// (longest surface sequence first, so that long tracks do not start last).
// Per outer iteration: idle lanes
// take new tracks; stepping lanes step until no lane steps any more, or at
// least `waitFrac`/64 of the busy lanes wait at a surface, or `maxInner`
// steps have been taken; then the waiting lanes run the surface update
// together, and finished lanes write their summary and go idle.  With Rev
// the kernel runs the reverse filter of the tracks that take one.  The
// navigator is a template parameter (Layers: the layer navigator): its
// option becomes a constant, so the direct navigator's instantiations carry
// none of the layer navigator's code (a run-time choice cost the direct
// fit's forward kernel 4 %).
template <bool Mixed, bool GeoInLds, bool Sympy = false, bool Rev = false, bool Layers = false>
__global__ __launch_bounds__(kBlock, KF_FORWARD_WAVES_PER_EU) void fitForwardKernel(
    FwdArgs args, int32_t* __restrict__ next, int32_t maxInner, int32_t waitFrac,
    int32_t quarantine, const int32_t* __restrict__ order) {
  FwdArgs a = args;
  a.o.navigator = Layers ? ACTS_KF_NAVIGATOR_LAYERS : ACTS_KF_NAVIGATOR_DIRECT;
  const int lane = static_cast<int>(threadIdx.x) & 63;
  // each lane's last field-map cell (cell-major maps): the first lookup of
  // an RK step is nearly always in the cell of the previous step's last one
  __shared__ double sCell[Mixed ? 1 : kCellValues * kCellCacheStride];
  __shared__ float sCellF[Mixed ? kCellValues * kCellCacheStride : 1];
  __shared__ int64_t sCellIdx[kCellCacheStride];
  static_assert(kBlock == kCellCacheStride, "one cache column per lane of the workgroup");
  a.F.cache = nullptr;
  a.F.cacheF = nullptr;
  a.F.cacheIdx = nullptr;
  if (Mixed ? a.F.cellsF != nullptr : a.F.cells != nullptr) {
    sCellIdx[lane] = -1;
    if constexpr (Mixed)
      a.F.cacheF = sCellF + lane;
    else
      a.F.cache = sCell + lane;
    a.F.cacheIdx = sCellIdx + lane;
  }
  const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t n = a.in.n;
  // Surfaces and material slabs are read at every step: with GeoInLds (the
  // geometry fits, as it does for the benchmark detectors) the workgroup
  // keeps a copy in LDS.  The choice is a template parameter so that every
  // geometry access compiles to an LDS read: a pointer that may point to
  // either memory needs flat loads, which wait for all outstanding global
  // loads and stores of the wavefront before their data can be used.
  __shared__ DevSurface sGeo[GeoInLds ? kLdsSurfaces : 1];
  __shared__ DevSlab sSlabs[GeoInLds ? kLdsSlabs : 1];
  if constexpr (GeoInLds) {
    const double* src = reinterpret_cast<const double*>(a.geo);
    double* dst = reinterpret_cast<double*>(sGeo);
    const int nd = a.nSurfaces * static_cast<int>(sizeof(DevSurface) / sizeof(double));
    for (int i = static_cast<int>(threadIdx.x); i < nd; i += blockDim.x) dst[i] = src[i];
    const int32_t* ss = reinterpret_cast<const int32_t*>(a.slabs);
    int32_t* sd_ = reinterpret_cast<int32_t*>(sSlabs);
    const int ns = a.nSlabs * static_cast<int>(sizeof(DevSlab) / sizeof(int32_t));
    for (int i = static_cast<int>(threadIdx.x); i < ns; i += blockDim.x) sd_[i] = ss[i];
    __syncthreads();
  }
  const DevSurface* geo = GeoInLds ? sGeo : a.geo;
  const DevSlab* slabs = GeoInLds ? sSlabs : a.slabs;
  Lane L;
  int phase = kPhaseIdle;
  while (true) {
    // a wavefront holding a track that has run longer than `quarantine`
    // steps takes no new tracks until it ends, so that the long track is not
    // slowed down by the surface updates of fresh ones
    const bool hold = quarantine > 0 &&
                      __ballot(phase == kPhaseStepping && L.iter > static_cast<uint32_t>(quarantine)) != 0;
    const uint64_t idle = hold ? 0ull : __ballot(phase == kPhaseIdle);
    if (idle) {
      const int cnt = __popcll(idle);
      const int leader = __ffsll(static_cast<long long>(idle)) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(next, cnt);
      base = __shfl(base, leader);
      if (phase == kPhaseIdle) {
        const int64_t q = static_cast<int64_t>(base) + __popcll(idle & lanesBelow);
        if (q >= n)
          phase = kPhaseRetired;
        else if constexpr (Rev)
          phase = laneInitReverse<Mixed, Sympy>(L, q, order ? order[q] : q, a, geo);
        else
          phase = laneInit<Mixed, Sympy>(L, q, order ? order[q] : q, a, geo);
      }
    }
    if (phase == kPhaseFinished) {
      laneFinish<Rev>(L, a);
      phase = kPhaseIdle;
    }
    const uint64_t busy = __ballot(phase != kPhaseRetired);
    if (!busy) break;
    for (int it = 0; it < maxInner; ++it) {
      if (phase == kPhaseStepping) phase = laneStep<Mixed, Sympy, Rev>(L, a, geo);
      const int nStep = __popcll(__ballot(phase == kPhaseStepping));
      const int nWait = __popcll(__ballot(phase == kPhaseSurface || phase == kPhaseFinished));
      if (nStep == 0 || 64 * nWait >= waitFrac * (nStep + nWait)) break;
    }
    // a lane that reached a surface while stepping without the transport
    // goes back to its checkpoint and repeats the steps with it; a track
    // that did so once (e.g. a looper that crosses a surface every turn)
    // keeps the transport from then on
    if (phase == kPhaseSurface && L.noTransport) {
      L = a.ckpt[static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x];
      L.replaying = true;
      L.noSkip = true;
      phase = kPhaseStepping;
    }
    if (phase == kPhaseSurface) phase = laneFilter<Mixed, Sympy, Rev>(L, a, geo, slabs);
    if (phase == kPhaseFinished) {
      laneFinish<Rev>(L, a);
      phase = kPhaseIdle;
    }
  }
}

// Track-state table: one thread per row
template <bool Sub>
__global__ __launch_bounds__(kTsBlock) void trackStatesKernel(TsArgs a) {
  const int64_t slot = static_cast<int64_t>(blockIdx.x) * kTsBlock + threadIdx.x;
  if (slot < a.nSlots) trackStateRow<Sub>(a, slot);
}

// Smoother (RTS, or MBF with Mbf) + track summary + outputs: one lane per
// track
template <bool Mbf>
__global__ __launch_bounds__(kBlock) void smoothKernel(DevBatch in, DevOut out,
                                                       DevScratch sc,
                                                       const int32_t* __restrict__ perm,
                                                       RevArgs rv, int32_t revEnabled) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool valid = g < in.n;
  const int64_t t = !valid ? 0 : (perm ? perm[g] : g);
  const bool rev = valid && revEnabled && reverseWanted(rv, sc, g);
  smoothTrack<Mbf>(g, t, valid, in, out, sc, rev);
}

// Extrapolation of every successfully smoothed track to the reference
// surface, one lane per track (KalmanFitter.hpp:959-971).  A failure fails
// the fit: the status becomes the extrapolation's error and the summary is
// cleared, as for any failed fit.
template <bool Mixed, bool Sympy = false>
__global__ __launch_bounds__(kBlock) void referenceKernel(DevOut out, DevScratch sc,
                                                          const DevSurface* __restrict__ geo,
                                                          DevField F, DevOpts o, LineRef L,
                                                          int32_t strategy, int64_t n,
                                                          const int64_t* __restrict__ seqOff,
                                                          const int32_t* __restrict__ perm) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  // failed fits, and reverse filters that reached the surface themselves
  if (q >= n || sc.track[q] != 0 || sc.track[2 * sc.stride + q] == kReferenceSet) return;
  const int64_t t = perm ? perm[q] : q;
  double p[6], C[36];
  const int st = extrapolateTrack<Mixed, Sympy>(q, sc, sc.ref, geo, F, o, L, strategy, p, C);
  if (st) {
    for (int64_t slot = seqOff[t]; slot < seqOff[t + 1]; ++slot) clearSlot(out, slot);
    if (out.status) out.status[t] = st;
    if (out.nStates) out.nStates[t] = 0;
    if (out.chi2) out.chi2[t] = 0.f;
    if (out.ndf) out.ndf[t] = 0;
    if (out.nMeas) out.nMeas[t] = 0;
    if (out.nHoles) out.nHoles[t] = 0;
    if (out.nOutliers) out.nOutliers[t] = 0;
    return;
  }
  if (out.referenceParams)
    for (int i = 0; i < 6; ++i) out.referenceParams[6 * t + i] = p[i];
  if (out.referenceCov)
    for (int i = 0; i < 36; ++i) out.referenceCov[36 * t + i] = C[i];
}

__global__ void zeroKernel(int32_t* p, int32_t count) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < count) p[i] = 0;
}

// Counting sort of the tracks by sequence length, longest first.
// Queue key of a track: its sequence length (longest tracks first), or
// maxLen + 1 for a track with a measurement within kSeamRisk rad of the +-pi
// seam of a cylinder.  There the measured loc0 = R phi and the prediction can
// lie on opposite sides of the seam; the residual is not wrapped (as in the
// reference), the update throws the track far off and it may run into the
// step limit, so such tracks are started first.  The key only orders the
// queue: results do not depend on it.
constexpr double kSeamRisk = 0.05;
struct QueueKeyArgs {
  const double* startParams;
  int32_t byDirection;  // 1: direction bins instead of the sequence length
  int32_t byMomentum;   // 1: sequence length, then momentum bins
  const int64_t* seqOff;
  const int64_t* measOff;
  const int32_t* measSurf;
  const double* measVal;
  const DevSurface* geo;
  int32_t nSurfaces;
  int32_t maxLen;
};
// With byDirection the key is instead a bin of the start direction and
// charge (16 eta bins over [-4, 4] x 32 phi bins x 2 charges, keys 0..1023,
// seam tracks 1024): neighbouring lanes then follow neighbouring paths, so
// that they reach their surfaces after similar numbers of steps and read
// the same cells of a field map.
constexpr int kDirectionKeys = 1024;
__device__ inline int32_t directionKey(const QueueKeyArgs& k, int64_t t) {
  const double* bp = k.startParams + 6 * t;
  const double theta = fmin(fmax(bp[3], 1e-6), kPi - 1e-6);
  const double eta = -log(tan(0.5 * theta));
  const int eb = min(max(static_cast<int>(floor((eta + 4.0) * 2.0)), 0), 15);
  const int pb = min(max(static_cast<int>(floor((bp[2] + kPi) * (16.0 / kPi))), 0), 31);
  const int qb = bp[4] < 0 ? 0 : 1;
  return (eb * 32 + pb) * 2 + qb;
}
// With byMomentum the key is the sequence length (capped at kMomentumLenCap)
// times kMomentumBins plus a bin of log |p| of the start parameters (0.1 to
// 1000 GeV): tracks of one length are ordered by momentum, so that the
// lanes of a wavefront bend alike and take similar numbers of steps between
// their surfaces.
constexpr int kMomentumBins = 16;
constexpr int kMomentumLenCap = 62;
__device__ inline int32_t momentumKey(const QueueKeyArgs& k, int64_t t, int32_t len) {
  const double p = 1.0 / fabs(k.startParams[6 * t + 4]);
  const int pb = min(max(static_cast<int>(floor((log10(p) + 1.0) * 4.0)), 0),
                     kMomentumBins - 1);
  return min(len, kMomentumLenCap) * kMomentumBins + pb;
}
// A track's queue key: the seam key if one of its measurements on a
// cylinder lies within kSeamRisk of phi = +-pi (such tracks may become
// loopers and go first), else its direction key, or its sequence length
// (capped at maxLen), or length and momentum (queueKeyKernel).
// Counting sort of the tracks by queue key, largest key first, in three
// passes without global atomics: per-block histograms (LDS atomics) stored
// key-major with the largest key first, so that one exclusive scan over the
// whole array gives every (key, block) its output offset; then a per-block
// scatter (LDS atomics for the rank inside a block and key; the order inside
// a bucket is not fixed, which only changes the queue order).
constexpr int kSortBlock = 256;
// tracks per workgroup of the histogram and scatter passes: the scanned
// array has nKeys entries per tile, so that with ~1 000 keys (direction or
// momentum orders) a 500 000-track batch scans 0.12 M entries, not the 2 M
// of one tile per 256 tracks (which took ~5 ms in one workgroup)
constexpr int kSortTile = 4096;
constexpr int kScanBlock = 1024;
constexpr int kMaxKeys = 1026;  // lengths above 1023 share a bucket

__device__ inline int64_t histIndex(int32_t key, int32_t nKeys, int64_t block, int64_t nBlocks) {
  return static_cast<int64_t>(nKeys - 1 - key) * nBlocks + block;
}

// The sort key of every track (its queue key, capped at kMaxKeys - 1),
// once: a workgroup takes
// kSortBlock consecutive tracks, reads their measurements with consecutive
// threads on consecutive measurements (a thread per track walking its own
// measurements touched a different cache line per lane and load and fetched
// ~7x the measurement arrays from HBM), marks in LDS the tracks with a
// cylinder measurement near the seam (the owner found by binary search
// over the workgroup's measurement offsets), then gives each track its key.
__global__ __launch_bounds__(kSortBlock) void queueKeyKernel(QueueKeyArgs k, int64_t n,
                                                             int32_t* keys) {
  __shared__ int64_t off[kSortBlock + 1];
  __shared__ int32_t seam[kSortBlock];
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kSortBlock;
  const int nt = static_cast<int>(min<int64_t>(kSortBlock, n - t0));
  for (int i = threadIdx.x; i <= nt; i += kSortBlock) off[i] = k.measOff[t0 + i];
  seam[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t m = off[0] + threadIdx.x; m < off[nt]; m += kSortBlock) {
    const int32_t s = k.measSurf[m];
    if (s < 0 || s >= k.nSurfaces || k.geo[s].type != ACTS_KF_SURFACE_CYLINDER) continue;
    const double phi = k.measVal[2 * m] / k.geo[s].b0;
    if (fabs(phi) > kPi - kSeamRisk) {
      int lo = 0, hi = nt;  // off[lo] <= m < off[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (off[mid] <= m)
          lo = mid;
        else
          hi = mid;
      }
      seam[lo] = 1;
    }
  }
  __syncthreads();
  if (threadIdx.x >= nt) return;
  const int64_t t = t0 + threadIdx.x;
  int32_t key;
  if (seam[threadIdx.x]) {
    key = k.byDirection ? kDirectionKeys
          : k.byMomentum ? (kMomentumLenCap + 1) * kMomentumBins
                         : k.maxLen + 1;
  } else if (k.byDirection) {
    key = directionKey(k, t);
  } else {
    const int32_t len = min(static_cast<int32_t>(k.seqOff[t + 1] - k.seqOff[t]), k.maxLen);
    key = k.byMomentum ? momentumKey(k, t, len) : len;
  }
  keys[t] = min(key, kMaxKeys - 1);
}

__global__ __launch_bounds__(kSortBlock) void keyHistogramKernel(const int32_t* keys, int64_t n,
                                                                 int32_t nKeys,
                                                                 int32_t* blockHist) {
  __shared__ int32_t h[kMaxKeys];
  for (int i = threadIdx.x; i < nKeys; i += kSortBlock) h[i] = 0;
  __syncthreads();
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
  const int64_t t1 = min<int64_t>(n, t0 + kSortTile);
  for (int64_t t = t0 + threadIdx.x; t < t1; t += kSortBlock) atomicAdd(&h[keys[t]], 1);
  __syncthreads();
  for (int i = threadIdx.x; i < nKeys; i += kSortBlock)
    blockHist[histIndex(i, nKeys, blockIdx.x, gridDim.x)] = h[i];
}

// Exclusive prefix sum of `count` ints in place, one workgroup: each thread
// sums a contiguous chunk, the chunk sums are scanned in LDS, each thread
// rewrites its chunk.
__global__ __launch_bounds__(kScanBlock) void exclusiveScanKernel(int32_t* a, int64_t count) {
  __shared__ int32_t part[kScanBlock];
  const int64_t chunk = (count + kScanBlock - 1) / kScanBlock;
  const int64_t lo = threadIdx.x * chunk;
  const int64_t hi = min(lo + chunk, count);
  int32_t sum = 0;
  for (int64_t i = lo; i < hi; ++i) sum += a[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t acc = 0;
    for (int i = 0; i < kScanBlock; ++i) {
      const int32_t c = part[i];
      part[i] = acc;
      acc += c;
    }
  }
  __syncthreads();
  int32_t acc = part[threadIdx.x];
  for (int64_t i = lo; i < hi; ++i) {
    const int32_t c = a[i];
    a[i] = acc;
    acc += c;
  }
}

// The scan of the whole histogram array in three passes over chunks of
// kScanChunk entries, so that many workgroups read it coalesced: the sum of
// every chunk (chunkSumKernel), the exclusive scan of those sums
// (exclusiveScanKernel, one workgroup over a few hundred entries), then each
// chunk scanned in LDS from its sum's offset (chunkScanKernel).
constexpr int kScanChunk = 4096;
constexpr int kScanThreads = 256;
__global__ __launch_bounds__(kScanThreads) void chunkSumKernel(const int32_t* a, int64_t count,
                                                               int32_t* sums) {
  __shared__ int32_t red[kScanThreads];
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * kScanChunk;
  const int64_t hi = min<int64_t>(count, lo + kScanChunk);
  int32_t sum = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kScanThreads) sum += a[i];
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int w = kScanThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kScanThreads) void chunkScanKernel(int32_t* a, int64_t count,
                                                                const int32_t* offsets) {
  constexpr int kPer = kScanChunk / kScanThreads;
  __shared__ int32_t v[kScanChunk];
  __shared__ int32_t part[kScanThreads];
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * kScanChunk;
  const int n = static_cast<int>(min<int64_t>(kScanChunk, count - lo));
  for (int i = threadIdx.x; i < n; i += kScanThreads) v[i] = a[lo + i];
  __syncthreads();
  const int first = static_cast<int>(threadIdx.x) * kPer;
  int32_t sum = 0;
  for (int j = 0; j < kPer; ++j)
    if (first + j < n) sum += v[first + j];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t acc = offsets[blockIdx.x];
    for (int t = 0; t < kScanThreads; ++t) {
      const int32_t c = part[t];
      part[t] = acc;
      acc += c;
    }
  }
  __syncthreads();
  int32_t acc = part[threadIdx.x];
  for (int j = 0; j < kPer; ++j) {
    if (first + j < n) {
      const int32_t c = v[first + j];
      v[first + j] = acc;
      acc += c;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kScanThreads) a[lo + i] = v[i];
}

__global__ __launch_bounds__(kSortBlock) void keyScatterKernel(const int32_t* keys, int64_t n,
                                                               int32_t nKeys,
                                                               const int32_t* blockOffset,
                                                               int32_t* perm) {
  __shared__ int32_t h[kMaxKeys];
  for (int i = threadIdx.x; i < nKeys; i += kSortBlock) h[i] = 0;
  __syncthreads();
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kSortTile;
  const int64_t t1 = min<int64_t>(n, t0 + kSortTile);
  for (int64_t t = t0 + threadIdx.x; t < t1; t += kSortBlock) {
    const int32_t key = keys[t];
    const int32_t rank = atomicAdd(&h[key], 1);
    perm[blockOffset[histIndex(key, nKeys, blockIdx.x, gridDim.x)] + rank] =
        static_cast<int32_t>(t);
  }
}

__global__ void maxLengthKernel(const int64_t* off, int64_t n, int32_t* result) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t len = static_cast<int32_t>(off[t + 1] - off[t]);
  atomicMax(result, len);
}

// Device buffer that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return 0;
    if (p) {
      if (hipFree(p) != hipSuccess) return -1;
      p = nullptr;
      bytes = 0;
    }
    if (need == 0) return 0;
    if (hipMalloc(&p, need) != hipSuccess) return -1;
    bytes = need;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// What a captured fit depends on: the caller's buffers, the scratch buffers
// and sizes, the stream.
struct GraphKey {
  acts_kf_batch_in in;
  acts_kf_batch_out out;
  int32_t states;
  int32_t pad;
  void* stream;
  void* scratch[5];
  void* ref;
};

PhysConst makePhysConst() {
  PhysConst c;
  const long double MeV = 1e-3L;
  c.me = static_cast<float>(static_cast<double>(1e-3 * 0.5109989461L));
  c.k = static_cast<float>(static_cast<double>(1e-3 * 30.7075L) * 1.0);
  c.plasmaScale = static_cast<float>(static_cast<double>(1e-9 * 28.816L));
  c.pad = 0.f;
  c.highland = static_cast<double>(1e-3 * 13.6L);
  c.rossiGreisen = static_cast<double>(1e-3 * 17.5L);
  c.minP = 10 * 1e-3;
  (void)MeV;
  return c;
}

}  // namespace

struct acts_kf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int32_t nSurfaces = 0;
  int32_t nSlabs = 0;
  DevSurface* geo = nullptr;
  DevSlab* slabs = nullptr;
  DevField field{};
  // field map: the point arrays (freed once the cell copies exist) and the
  // cell-major copies the kernels read
  double* fieldValues = nullptr;
  float* fieldValuesF = nullptr;
  double* fieldCells = nullptr;
  float* fieldCellsF = nullptr;
  bool mixed = false;
  // kernel timing: events before the forward kernel, between the two
  // kernels and after the smoother
  bool timing = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  bool timed = false;
  // recorded after the last work of every call on whichever stream it used:
  // acts_kf_sync waits for it and for nothing else on the device
  hipEvent_t done = nullptr;
  bool pending = false;
  // captured launch sequence of the last fit, replayed while its key holds
  bool useGraph = false;
  hipGraphExec_t graphExec = nullptr;
  GraphKey graphKey{};
  DevOpts opts{};
  PhysConst pc{};
  DevBuf scratchD, scratchI, scratchT, lenBuf, permBuf, histBuf, queueBuf;
  // track queue order: 0 input order, 1 longest sequence first, 2 by start
  // direction (kf: directionKey)
  int32_t queueOrder = 1;
  // persistent forward kernel: resident wavefronts and scheduling knobs
  int64_t residentWaves = 0;
  int32_t innerSteps = 64;
  int32_t waitFrac = 64;
  int32_t quarantine = 0;
  bool mbf = false;  // MbfSmoother instead of GainMatrixSmoother
  bool sympy = false;  // SympyStepper instead of EigenStepper
  // lanes stepping without the transport save their state here, one Lane
  // per resident lane of the forward kernel
  int32_t skipAfter = 0;
  DevBuf ckptBuf;
  // reverse filtering (KalmanFitterOptions::reverseFiltering, the covariance
  // scaling, the reverse-filtering logic's momentum threshold)
  bool revAll = false;
  bool revEnabled = false;
  double revScale = 100.0;
  double revThreshold = 0.0;
  DevBuf revBuf;
  // reference surface of the fitted parameters, and the default propagator
  // and stepper options its extrapolation runs with
  int32_t refSurface = ACTS_KF_REFERENCE_NONE;
  int32_t refStrategy = ACTS_KF_EXTRAPOLATE_FIRST_OR_LAST;
  LineRef refLine{};
  DevOpts extOpts{};
  DevBuf refBuf;
  // staging for the host-buffer entry point
  DevBuf inBuf, outBuf;
  // device address of the static workspace on this context's device
  void* staticWs = nullptr;
};

namespace {

// Cell-major copy of a point-array field map (DevField::cells): one thread
// per cell gathers its 8 corners (a z pair for each (x, y) corner) and
// writes them as one record of `Stride` values, the tail zero.
template <typename T, int Stride>
__global__ void fieldCellsKernel(const T* __restrict__ v, T* __restrict__ cells, int32_t n0,
                                 int32_t n1, int32_t n2) {
  const int64_t cell = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (cell >= static_cast<int64_t>(n0) * n1 * n2) return;
  const int64_t iz = cell % n2;
  const int64_t ixy = cell / n2;
  const int64_t iy = ixy % n1;
  const int64_t ix = ixy / n1;
  const int64_t sy = n2 + 1;
  const int64_t sx = static_cast<int64_t>(n1 + 1) * (n2 + 1);
  T* out = cells + static_cast<int64_t>(Stride) * cell;
  for (int i = 0; i < 4; ++i)
    for (int dz = 0; dz < 2; ++dz)
      for (int k = 0; k < 3; ++k)
        out[6 * i + 3 * dz + k] = v[3 * ((ix + (i >> 1)) * sx + (iy + (i & 1)) * sy + iz + dz) + k];
  for (int j = kCellValues; j < Stride; ++j) out[j] = T(0);
}

// Device address of `scratch` on device `device` (current), fetched once
// per device: each device holds its own copy of the array.
std::mutex g_staticWsMutex;
void* g_staticWs[kMaxDevices] = {};

int staticWorkspace(int device, void** out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(ACTS_KF_ERR_UNSUPPORTED, "device index beyond the static workspace table");
  std::lock_guard<std::mutex> lock(g_staticWsMutex);
  if (!g_staticWs[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ACTS_KF_ERR_DEVICE,
                  std::string("hipGetSymbolAddress(static workspace): ") + hipGetErrorString(e));
    g_staticWs[device] = p;
  }
  *out = g_staticWs[device];
  return ACTS_KF_OK;
}

// Field maps above this size stay in the point layout on the device.
constexpr size_t kMaxCellBytes = size_t{32} << 30;

// Uploads the padded point array of a field map (double, or float for the
// mixed precision) and expands it into the cell-major layout the kernels
// read; if the cell copy does not fit, the kernels read the point array.
template <typename T, int Stride>
int uploadPoints(acts_kf_ctx* c, const std::vector<double>& padded, T** pts, T** cells) {
  std::vector<T> hv(padded.begin(), padded.end());
  if (hipMalloc(pts, sizeof(T) * hv.size()) != hipSuccess)
    return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "hipMalloc field map failed");
  if (hipMemcpy(*pts, hv.data(), sizeof(T) * hv.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(ACTS_KF_ERR_DEVICE, "hipMemcpy field map failed");
  const DevField& F = c->field;
  const int64_t nCells = static_cast<int64_t>(F.nb[0]) * F.nb[1] * F.nb[2];
  const size_t cellBytes = sizeof(T) * Stride * static_cast<size_t>(nCells);
  if (cellBytes > kMaxCellBytes) return ACTS_KF_OK;
  if (hipMalloc(cells, cellBytes) != hipSuccess) {
    (void)hipGetLastError();  // not fatal: the point layout is used
    *cells = nullptr;
    return ACTS_KF_OK;
  }
  const unsigned blocks = static_cast<unsigned>((nCells + 255) / 256);
  hipLaunchKernelGGL((fieldCellsKernel<T, Stride>), dim3(blocks), dim3(256), 0, c->stream, *pts,
                     *cells, F.nb[0], F.nb[1], F.nb[2]);
  KF_HIP(hipGetLastError());
  KF_HIP(hipStreamSynchronize(c->stream));
  (void)hipFree(*pts);
  *pts = nullptr;
  return ACTS_KF_OK;
}

int uploadFieldMap(acts_kf_ctx* c, const std::vector<double>& padded) {
  if (!c->mixed) {
    const int r = uploadPoints<double, kCellDoubles>(c, padded, &c->fieldValues, &c->fieldCells);
    c->field.v = c->fieldValues;
    c->field.cells = c->fieldCells;
    return r;
  }
  // the mixed-precision fit reads an FP32 copy of the map
  const int r = uploadPoints<float, kCellFloats>(c, padded, &c->fieldValuesF, &c->fieldCellsF);
  c->field.vf = c->fieldValuesF;
  c->field.cellsF = c->fieldCellsF;
  return r;
}

using ForwardFn = void (*)(FwdArgs, int32_t*, int32_t, int32_t, int32_t, const int32_t*);

bool geometryFitsLds(const acts_kf_ctx* c) {
  return c->nSurfaces <= kLdsSurfaces && c->nSlabs <= kLdsSlabs;
}

template <bool Rev>
ForwardFn forwardKernelOf(bool mixed, bool geoInLds, bool sympy, bool layers) {
  if (layers) {
    if (sympy && mixed)
      return geoInLds ? fitForwardKernel<true, true, true, Rev, true>
                      : fitForwardKernel<true, false, true, Rev, true>;
    if (sympy)
      return geoInLds ? fitForwardKernel<false, true, true, Rev, true>
                      : fitForwardKernel<false, false, true, Rev, true>;
    if (mixed)
      return geoInLds ? fitForwardKernel<true, true, false, Rev, true>
                      : fitForwardKernel<true, false, false, Rev, true>;
    return geoInLds ? fitForwardKernel<false, true, false, Rev, true>
                    : fitForwardKernel<false, false, false, Rev, true>;
  }
  if (sympy && mixed)
    return geoInLds ? fitForwardKernel<true, true, true, Rev>
                    : fitForwardKernel<true, false, true, Rev>;
  if (sympy)
    return geoInLds ? fitForwardKernel<false, true, true, Rev>
                    : fitForwardKernel<false, false, true, Rev>;
  if (mixed)
    return geoInLds ? fitForwardKernel<true, true, false, Rev>
                    : fitForwardKernel<true, false, false, Rev>;
  return geoInLds ? fitForwardKernel<false, true, false, Rev>
                  : fitForwardKernel<false, false, false, Rev>;
}
// the forward filter, or with `rev` the reverse filter, of a context
ForwardFn forwardKernel(const acts_kf_ctx* c, bool rev = false) {
  const bool lds = geometryFitsLds(c);
  const bool layers = c->opts.navigator != ACTS_KF_NAVIGATOR_DIRECT;
  return rev ? forwardKernelOf<true>(c->mixed, lds, c->sympy, layers)
             : forwardKernelOf<false>(c->mixed, lds, c->sympy, layers);
}

}  // namespace

extern "C" {

void acts_kf_default_options(acts_kf_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->multiple_scattering = 1;
  o->energy_loss = 1;
  o->use_joseph = 0;
  o->precision = ACTS_KF_PRECISION_FP64;
  o->outlier_chi2_cut = 0.0;
  o->step_tolerance = 1e-4;
  o->step_size_cutoff = 0.0;
  o->initial_step_size = 10000.0;
  o->max_step_size = DBL_MAX;
  o->max_runge_kutta_step_trials = 10000;
  o->max_steps = 1000;
  o->max_target_skipping = 100;
  o->loop_protection = 1;
  o->path_limit = DBL_MAX;
  o->loop_fraction = 0.5;
  o->surface_tolerance = 1e-4;
  o->near_limit = -100 * 1e-3;
  o->abs_pdg = 13;
  o->mass = 0.1056583755f;
  o->abs_charge = 1.0f;
  o->length_sorted = 1;
  o->reverse_filtering = 0;
  o->reverse_filtering_cov_scaling = 100.0;
  o->reverse_filtering_momentum_threshold = 0.0;
}

const char* acts_kf_last_error(void) { return g_lastError.c_str(); }

const char* acts_kf_status_string(int32_t s) {
  switch (s) {
    case ACTS_KF_STATUS_OK: return "success";
    case ACTS_KF_STATUS_UPDATE_FAILED: return "Kalman update failed";
    case ACTS_KF_STATUS_SMOOTH_FAILED: return "Kalman smooth failed";
    case ACTS_KF_CAT_KALMAN | 3: return "Kalman output conversion failed";
    case ACTS_KF_STATUS_NO_MEASUREMENT_FOUND:
      return "No measurement detected during the propagation";
    case ACTS_KF_CAT_KALMAN | 5: return "Reverse propagation failed";
    case ACTS_KF_CAT_KALMAN | 6:
      return "Inconsistent reference surfaces between forward and reversed filtered tracks";
    case ACTS_KF_STATUS_PROPAGATION_FAILURE: return "Propagation failed";
    case ACTS_KF_STATUS_STEP_COUNT_LIMIT:
      return "Propagation reached the configured maximum number of steps";
    case ACTS_KF_STATUS_NEXT_TARGET_LIMIT:
      return "Propagation reached the configured maximum number of next target calls";
    case ACTS_KF_STATUS_STEP_SIZE_STALLED: return "Step size fell below minimum threshold";
    case ACTS_KF_STATUS_STEP_INVALID: return "Step was invalid";
    case ACTS_KF_STATUS_STEP_ADJUSTMENT_FAILED:
      return "Step size adjustment exceeds maximum trials";
    case ACTS_KF_STATUS_NOT_ON_SURFACE:
      return "Global to local transformation failed: position not on surface.";
    case ACTS_KF_STATUS_FIELD_OUT_OF_BOUNDS:
      return "The lookup position was outside of the validity domain of the field";
    case ACTS_KF_STATUS_NO_COMPATIBLE_STATE: return "No compatible track state found";
    case ACTS_KF_STATUS_REFERENCE_UNREACHABLE: return "Reference surface is unreachable";
    case ACTS_KF_STATUS_STATE_CAPACITY: return "More track states than the track has slots";
    default: return "unknown";
  }
}

int acts_kf_create(const acts_kf_surface* surfaces, int32_t nSurfaces,
                   const acts_kf_material_slab* materials, int32_t nMaterials,
                   const acts_kf_field* field, const acts_kf_options* options,
                   int32_t device, acts_kf_ctx** outCtx) {
  if (!outCtx || !surfaces || nSurfaces <= 0 || !field || !options)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: null or empty argument");
  if (nMaterials > 0 && !materials)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: materials is null");
  if (field->kind == ACTS_KF_FIELD_GRID_XYZ) {
    if (!field->values)
      return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: field grid values are null");
    for (int a = 0; a < 3; ++a)
      if (field->n[a] < 2 || field->n[a] > (1 << 20) ||
          !(field->point_max[a] > field->point_min[a]))
        return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: bad field grid axis");
  } else if (field->kind != ACTS_KF_FIELD_CONSTANT) {
    return fail(ACTS_KF_ERR_UNSUPPORTED, "acts_kf_create: field kind not supported");
  }
  if (options->precision != ACTS_KF_PRECISION_FP64 &&
      options->precision != ACTS_KF_PRECISION_MIXED)
    return fail(ACTS_KF_ERR_UNSUPPORTED, "acts_kf_create: precision mode not supported");
  if (options->stepper != ACTS_KF_STEPPER_EIGEN && options->stepper != ACTS_KF_STEPPER_SYMPY)
    return fail(ACTS_KF_ERR_UNSUPPORTED, "acts_kf_create: stepper not supported");
  if (options->smoother != ACTS_KF_SMOOTHER_GAIN_MATRIX &&
      options->smoother != ACTS_KF_SMOOTHER_MBF)
    return fail(ACTS_KF_ERR_UNSUPPORTED, "acts_kf_create: smoother not supported");
  if (options->reference_surface != ACTS_KF_REFERENCE_NONE &&
      options->reference_surface != ACTS_KF_REFERENCE_PERIGEE)
    return fail(ACTS_KF_ERR_UNSUPPORTED, "acts_kf_create: reference surface not supported");
  if (options->reference_strategy < ACTS_KF_EXTRAPOLATE_FIRST ||
      options->reference_strategy > ACTS_KF_EXTRAPOLATE_FIRST_OR_LAST)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: bad extrapolation strategy");
  for (int32_t i = 0; i < nSurfaces; ++i) {
    const acts_kf_surface& s = surfaces[i];
    if (s.type < 0 || s.type > 2)
      return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: bad surface type");
    if (s.material >= nMaterials)
      return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: bad material index");
  }
  int nDev = 0;
  KF_HIP(hipGetDeviceCount(&nDev));
  if (device < 0 || device >= nDev)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: bad device index");
  KF_HIP(hipSetDevice(device));

  std::vector<DevSurface> hs(nSurfaces);
  for (int32_t i = 0; i < nSurfaces; ++i) {
    const acts_kf_surface& s = surfaces[i];
    DevSurface& d = hs[i];
    std::memcpy(d.R, s.rotation, sizeof(d.R));
    std::memcpy(d.t, s.translation, sizeof(d.t));
    invert3(d.R, d.Ri);
    for (int r = 0; r < 3; ++r)
      d.ti[r] = -((d.Ri[3 * r] * d.t[0] + d.Ri[3 * r + 1] * d.t[1]) + d.Ri[3 * r + 2] * d.t[2]);
    d.b0 = s.bounds[0];
    d.b1 = s.bounds[1];
    d.type = s.type;
    d.flags = static_cast<int32_t>(s.flags);
    d.mat = s.material;
    d.pad = 0;
  }
  std::vector<DevSlab> hm(std::max(nMaterials, 1));
  for (int32_t i = 0; i < nMaterials; ++i) {
    const acts_kf_material_slab& m = materials[i];
    DevSlab& d = hm[i];
    // Material::fromMolarDensity + MaterialSlab (Material.cpp, MaterialSlab.hpp)
    const float ne = m.z * m.molar_rho;
    const float meanI =
        static_cast<float>(static_cast<double>(1e-9 * 16.0L) * std::pow(m.z, 0.9f));
    const float eps = 2 * FLT_EPSILON;
    d.thickness = m.thickness;
    d.tInX0 = (eps < m.x0) ? (m.thickness / m.x0) : 0;
    d.tInL0 = (eps < m.l0) ? (m.thickness / m.l0) : 0;
    d.molarElectronRho = ne;
    d.meanI = meanI;
    d.vacuum = (m.ar <= 0.f) ? 1 : 0;
  }

  acts_kf_ctx* c = new acts_kf_ctx();
  c->device = device;
  c->nSurfaces = nSurfaces;
  auto cleanup = [&](int code, const std::string& msg) {
    acts_kf_destroy(c);
    return fail(code, msg);
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(ACTS_KF_ERR_DEVICE, "hipStreamCreate failed");
  if (hipEventCreateWithFlags(&c->done, hipEventDisableTiming) != hipSuccess)
    return cleanup(ACTS_KF_ERR_DEVICE, "hipEventCreate failed");
  if (const int r = staticWorkspace(device, &c->staticWs)) {
    acts_kf_destroy(c);
    return r;
  }
  if (hipMalloc(&c->geo, sizeof(DevSurface) * nSurfaces) != hipSuccess ||
      hipMalloc(&c->slabs, sizeof(DevSlab) * hm.size()) != hipSuccess)
    return cleanup(ACTS_KF_ERR_OUT_OF_MEMORY, "hipMalloc geometry failed");
  c->nSlabs = static_cast<int32_t>(hm.size());
  if (hipMemcpy(c->geo, hs.data(), sizeof(DevSurface) * nSurfaces, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(c->slabs, hm.data(), sizeof(DevSlab) * hm.size(), hipMemcpyHostToDevice) !=
          hipSuccess)
    return cleanup(ACTS_KF_ERR_DEVICE, "hipMemcpy geometry failed");
  c->field.kind = field->kind;
  for (int i = 0; i < 3; ++i) c->field.B[i] = field->constant[i];
  c->mixed = options->precision == ACTS_KF_PRECISION_MIXED;
  if (field->kind == ACTS_KF_FIELD_GRID_XYZ) {
    std::vector<double> padded;
    if (!buildGridField(*field, c->field, padded))
      return cleanup(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: bad field grid");
    if (c->mixed) {
      for (int a = 0; a < 3; ++a) {
        c->field.minF[a] = static_cast<float>(c->field.min[a]);
        c->field.widthF[a] = static_cast<float>(c->field.width[a]);
        c->field.upperF[a] = static_cast<float>(c->field.upper[a]);
      }
    }
    const int r = uploadFieldMap(c, padded);
    if (r) {
      acts_kf_destroy(c);
      return r;
    }
  }
  DevOpts& o = c->opts;
  o.stepTolerance = options->step_tolerance;
  o.stepSizeCutOff = options->step_size_cutoff;
  o.initialStepSize = options->initial_step_size;
  o.maxStepSize = options->max_step_size;
  o.pathLimit = options->path_limit;
  o.loopFraction = options->loop_fraction;
  o.surfaceTolerance = options->surface_tolerance;
  o.nearLimit = options->near_limit;
  o.chi2Cut = options->outlier_chi2_cut;
  o.maxTrials = options->max_runge_kutta_step_trials;
  o.maxSteps = options->max_steps;
  o.maxTargetSkipping = options->max_target_skipping;
  o.loopProtection = options->loop_protection;
  o.multipleScattering = options->multiple_scattering;
  o.energyLoss = options->energy_loss;
  o.useJoseph = options->use_joseph;
  o.absPdg = options->abs_pdg;
  o.mass = options->mass;
  o.absQ = options->abs_charge;
  if (options->navigator != ACTS_KF_NAVIGATOR_DIRECT &&
      options->navigator != ACTS_KF_NAVIGATOR_LAYERS) {
    acts_kf_destroy(c);
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: bad navigator");
  }
  o.navigator = options->navigator;
  c->pc = makePhysConst();
  c->mbf = options->smoother == ACTS_KF_SMOOTHER_MBF;
  c->sympy = options->stepper == ACTS_KF_STEPPER_SYMPY;
  c->refSurface = options->reference_surface;
  c->refStrategy = options->reference_strategy;
  for (int i = 0; i < 3; ++i) c->refLine.c[i] = options->reference_position[i];
  // extrapolateTrackToReferenceSurface runs with default-constructed
  // propagator options (KalmanFitter.hpp:960-961): the defaults of
  // PropagatorOptions.hpp:26-52 and StepperOptions.hpp:36-49, the fit's
  // particle hypothesis
  {
    acts_kf_options d;
    acts_kf_default_options(&d);
    DevOpts& e = c->extOpts;
    e = o;
    e.stepTolerance = d.step_tolerance;
    e.stepSizeCutOff = d.step_size_cutoff;
    e.initialStepSize = d.initial_step_size;
    e.maxStepSize = d.max_step_size;
    e.pathLimit = d.path_limit;
    e.loopFraction = d.loop_fraction;
    e.surfaceTolerance = d.surface_tolerance;
    e.nearLimit = d.near_limit;
    e.maxTrials = d.max_runge_kutta_step_trials;
    e.maxSteps = d.max_steps;
    e.maxTargetSkipping = d.max_target_skipping;
    e.loopProtection = d.loop_protection;
  }
  if (options->length_sorted < 0 || options->length_sorted > 3)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_create: bad length_sorted (queue order)");
  c->queueOrder = options->length_sorted;
  c->timing = options->kernel_timing != 0;
  c->useGraph = options->use_graph != 0;
  if (c->timing)
    for (auto& e : c->ev)
      if (hipEventCreate(&e) != hipSuccess) {
        acts_kf_destroy(c);
        return fail(ACTS_KF_ERR_DEVICE, "hipEventCreate failed");
      }
  if (options->sched_inner_steps > 0) c->innerSteps = options->sched_inner_steps;
  if (options->sched_wait_64ths > 0) c->waitFrac = std::min(options->sched_wait_64ths, 64);
  c->quarantine = std::max(options->sched_quarantine_steps, 0);
  // the layer navigator keeps its resolved list in device memory, which a
  // transport-skip checkpoint would not restore: no skipping with it
  c->skipAfter = o.navigator ? 0 : std::max(options->sched_transport_skip_steps, 0);
  c->revAll = options->reverse_filtering != 0;
  c->revScale = options->reverse_filtering_cov_scaling;
  c->revThreshold = options->reverse_filtering_momentum_threshold;
  c->revEnabled = c->revAll || c->revThreshold > 0;
  {
    int blocksPerCu = 0, cus = 0;
    const hipError_t oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocksPerCu, forwardKernel(c), kBlock, 0);
    if (oe != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        blocksPerCu <= 0 || cus <= 0) {
      acts_kf_destroy(c);
      return fail(ACTS_KF_ERR_DEVICE, "occupancy query failed");
    }
    c->residentWaves = static_cast<int64_t>(blocksPerCu) * cus;
    if (c->skipAfter > 0 &&
        c->ckptBuf.ensure(static_cast<size_t>(c->residentWaves) * kBlock * sizeof(Lane))) {
      acts_kf_destroy(c);
      return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "acts_kf_create: checkpoint buffer");
    }
  }
  *outCtx = c;
  return ACTS_KF_OK;
}

// Records the context's completion event on `s` after the work just queued.
static int markQueued(acts_kf_ctx* c, hipStream_t s) {
  KF_HIP(hipEventRecord(c->done, s));
  c->pending = true;
  return ACTS_KF_OK;
}

// Offsets of a host-side batch: both CSR offset arrays start at 0 and never
// decrease (anything else would index outside the input and output arrays).
static bool offsetsValid(const int64_t* off, int64_t n) {
  if (off[0] != 0) return false;
  for (int64_t t = 0; t < n; ++t)
    if (off[t + 1] < off[t]) return false;
  return true;
}

// Device memory a fit works in besides its inputs and outputs.
struct Workspace {
  double* d = nullptr;       // per-state records (DevScratch::d)
  int32_t* i = nullptr;      // their int parts (DevScratch::i)
  int32_t* track = nullptr;  // per track: status, tip, first/last measurement
  double* ref = nullptr;     // with a reference surface: first/last smoothed states
  int32_t* queue = nullptr;  // the forward kernel's track-queue counter
  int32_t* perm = nullptr;   // queue order (sorted queues only)
  int32_t* hist = nullptr;   // per-block key histograms of the sort
  Lane* ckpt = nullptr;      // transport-skip checkpoints (null: no skipping)
  double* rev = nullptr;     // reverse filtering: per slot its record (kRevRecord)
  int32_t* revReached = nullptr;  // and whether it reached the reference surface
};

// Sizes of the workspace of a fit of n tracks of at most `states` states.
struct WorkspaceSizes {
  int32_t states = 1;
  int64_t stride = 0;  // slots per record component: n rounded up to a wavefront
  bool withRef = false, sorted = false, byDirection = false, byMomentum = false;
  int32_t nKeys = 0;
  int64_t sortBlocks = 0;
  size_t d = 0, i = 0, track = 0, ref = 0, queue = 0, perm = 0, hist = 0;
  size_t rev = 0;  // reverse filtering: records and reached flags
};

// 64-bit words of the layer navigator's reached-layer mask per track.
static int64_t navMaskWords(const acts_kf_ctx* c) { return (c->nSurfaces + 63) / 64; }

static WorkspaceSizes workspaceSizes(const acts_kf_ctx* c, int64_t n, int32_t maxLen) {
  WorkspaceSizes w;
  w.states = std::max(maxLen, 1);
  w.stride = (n + kBlock - 1) / kBlock * kBlock;
  w.d = sizeof(double) * static_cast<size_t>(w.states) * kScratchComps * w.stride;
  w.i = sizeof(int32_t) * static_cast<size_t>(w.states) * kIScratchComps * w.stride;
  // per track: status, tip, first and last measurement; the layer
  // navigator's record after them (kNavComps doubles per slot, then one
  // 64-bit word of its reached-layer mask per 64 surfaces)
  w.track = sizeof(int32_t) * 4 * w.stride +
            (c->opts.navigator ? sizeof(double) * (kNavComps + navMaskWords(c)) * w.stride : 0);
  w.withRef = c->refSurface != ACTS_KF_REFERENCE_NONE;
  w.ref = w.withRef ? sizeof(double) * kRefComps * static_cast<size_t>(n) : 0;
  w.queue = sizeof(int32_t);
  w.sorted = c->queueOrder != 0 && n > kBlock;
  w.byDirection = c->queueOrder == 2;
  w.byMomentum = c->queueOrder == 3;
  w.nKeys = w.byDirection  ? kDirectionKeys + 1
            : w.byMomentum ? (kMomentumLenCap + 1) * kMomentumBins + 1
                           : std::min(w.states + 2, kMaxKeys);
  w.sortBlocks = (n + kSortTile - 1) / kSortTile;
  // the queue permutation, then the tracks' sort keys
  w.perm = w.sorted ? sizeof(int32_t) * 2 * n : 0;
  // the histogram, then one sum per scan chunk
  w.hist = w.sorted ? sizeof(int32_t) * (w.nKeys * w.sortBlocks +
                                         (w.nKeys * w.sortBlocks + kScanChunk - 1) / kScanChunk)
                    : 0;
  w.rev = c->revEnabled
              ? (sizeof(double) * kRevRecord + sizeof(int32_t)) * static_cast<size_t>(n) + 256
              : 0;
  return w;
}

// The reverse filter's records and reached flags in a block of w.rev bytes.
static void reverseSections(char* base, const WorkspaceSizes& w, Workspace& ws) {
  const size_t n = (w.rev - 256) / (sizeof(double) * kRevRecord + sizeof(int32_t));
  ws.rev = reinterpret_cast<double*>(base);
  ws.revReached = reinterpret_cast<int32_t*>(base + sizeof(double) * kRevRecord * n);
}

// The context's own (grow-only) workspace buffers, grown to `w`.
static int contextWorkspace(acts_kf_ctx* c, const WorkspaceSizes& w, Workspace& ws) {
  if (w.withRef && c->refBuf.ensure(w.ref))
    return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "reference scratch allocation failed");
  if (c->scratchD.ensure(w.d) || c->scratchI.ensure(w.i) || c->scratchT.ensure(w.track) ||
      c->queueBuf.ensure(w.queue))
    return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "scratch allocation failed");
  if (w.sorted && (c->permBuf.ensure(w.perm) || c->histBuf.ensure(w.hist)))
    return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "permutation allocation failed");
  if (w.rev && c->revBuf.ensure(w.rev))
    return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "reverse-filter scratch allocation failed");
  if (w.rev) reverseSections(static_cast<char*>(c->revBuf.p), w, ws);
  ws.d = static_cast<double*>(c->scratchD.p);
  ws.i = static_cast<int32_t*>(c->scratchI.p);
  ws.track = static_cast<int32_t*>(c->scratchT.p);
  ws.ref = w.withRef ? static_cast<double*>(c->refBuf.p) : nullptr;
  ws.queue = static_cast<int32_t*>(c->queueBuf.p);
  ws.perm = w.sorted ? static_cast<int32_t*>(c->permBuf.p) : nullptr;
  ws.hist = w.sorted ? static_cast<int32_t*>(c->histBuf.p) : nullptr;
  ws.ckpt = c->skipAfter > 0 ? static_cast<Lane*>(c->ckptBuf.p) : nullptr;
  return ACTS_KF_OK;
}

// Queues the fit of a batch on `stream` in the workspace `ws` (sized by
// `w`); with `graph` the launch sequence is captured and replayed while
// nothing it depends on changes.
static int launchFit(acts_kf_ctx* c, const acts_kf_batch_in* in,
                     const acts_kf_batch_out* out, hipStream_t stream,
                     const WorkspaceSizes& w, const Workspace& ws, bool graph) {
  const int64_t n = in->n_tracks;
  if (n == 0) return ACTS_KF_OK;
  if (n > kMaxBatch) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "batch too large");
  const int64_t stride = w.stride;
  const int32_t states = w.states;
  const bool withRef = w.withRef;
  const bool sorted = w.sorted;
  const bool byDirection = w.byDirection;
  const int32_t nKeys = w.nKeys;
  const int64_t sortBlocks = w.sortBlocks;
  // Replay the captured launch sequence when nothing it depends on changed.
  GraphKey key;
  if (graph) {
    std::memset(&key, 0, sizeof(key));
    key.in = *in;
    key.out = *out;
    key.states = states;
    key.stream = stream;
    key.scratch[0] = ws.d;
    key.scratch[1] = ws.i;
    key.scratch[2] = ws.track;
    key.scratch[3] = ws.perm;
    key.scratch[4] = ws.hist;
    key.ref = ws.ref;
    if (c->graphExec && std::memcmp(&key, &c->graphKey, sizeof(key)) == 0) {
      KF_HIP(hipGraphLaunch(c->graphExec, stream));
      return ACTS_KF_OK;
    }
    if (c->graphExec) {
      (void)hipGraphExecDestroy(c->graphExec);
      c->graphExec = nullptr;
    }
    KF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  }
  const bool timing = c->timing && !graph;
  DevBatch b;
  b.n = n;
  b.startParams = in->start_params;
  b.startCov = in->start_cov;
  b.startPlane = in->start_plane;
  b.seqOff = in->seq_offsets;
  b.seqSurf = in->seq_surfaces;
  b.measOff = in->meas_offsets;
  b.measSurf = in->meas_surfaces;
  b.measVal = in->meas_values;
  b.measCov = in->meas_cov;
  b.measDims = in->meas_dims;
  b.measIdx = in->meas_indices;
  DevOut d;
  d.status = out->status;
  d.nStates = out->n_states;
  d.chi2 = out->chi2;
  d.ndf = out->ndf;
  d.nMeas = out->n_measurements;
  d.nHoles = out->n_holes;
  d.nOutliers = out->n_outliers;
  d.nSteps = out->n_steps;
  d.nTrials = out->n_step_trials;
  d.stateSurface = out->state_surface;
  d.stateFlags = out->state_flags;
  d.statePath = out->state_path;
  d.stateChi2 = out->state_chi2;
  d.predicted = out->predicted;
  d.predictedCov = out->predicted_cov;
  d.filtered = out->filtered;
  d.filteredCov = out->filtered_cov;
  d.smoothed = out->smoothed;
  d.smoothedCov = out->smoothed_cov;
  d.jacobian = out->jacobian;
  d.referenceParams = out->reference_params;
  d.referenceCov = out->reference_cov;
  d.stateMeas = out->state_measurement;
  d.reversed = out->reversed;
  DevScratch sc;
  sc.ref = withRef ? ws.ref : nullptr;
  sc.d = ws.d;
  sc.i = ws.i;
  sc.track = ws.track;
  sc.nav = c->opts.navigator ? reinterpret_cast<double*>(ws.track + 4 * stride) : nullptr;
  sc.navMask = sc.nav ? reinterpret_cast<uint64_t*>(sc.nav + kNavComps * stride) : nullptr;
  sc.stride = stride;
  sc.maxStates = states;
  const int64_t blocks = (n + kBlock - 1) / kBlock;
  int32_t* perm = nullptr;
  if (sorted) {
    perm = ws.perm;
    int32_t* hist = ws.hist;
    QueueKeyArgs qk;
    qk.startParams = in->start_params;
    qk.byDirection = byDirection ? 1 : 0;
    qk.byMomentum = w.byMomentum ? 1 : 0;
    qk.seqOff = in->seq_offsets;
    qk.measOff = in->meas_offsets;
    qk.measSurf = in->meas_surfaces;
    qk.measVal = in->meas_values;
    qk.geo = c->geo;
    qk.nSurfaces = c->nSurfaces;
    qk.maxLen = states;
    const unsigned nb = static_cast<unsigned>(sortBlocks);
    const unsigned nbKey = static_cast<unsigned>((n + kSortBlock - 1) / kSortBlock);
    // the keys after the permutation in the same buffer (workspaceSizes)
    int32_t* keys = perm + n;
    hipLaunchKernelGGL(queueKeyKernel, dim3(nbKey), dim3(kSortBlock), 0, stream, qk, n, keys);
    hipLaunchKernelGGL(keyHistogramKernel, dim3(nb), dim3(kSortBlock), 0, stream, keys, n,
                       nKeys, hist);
    const int64_t count = static_cast<int64_t>(nKeys) * sortBlocks;
    const int64_t chunks = (count + kScanChunk - 1) / kScanChunk;
    int32_t* sums = hist + count;
    hipLaunchKernelGGL(chunkSumKernel, dim3(static_cast<unsigned>(chunks)), dim3(kScanThreads), 0,
                       stream, hist, count, sums);
    hipLaunchKernelGGL(exclusiveScanKernel, dim3(1), dim3(kScanBlock), 0, stream, sums, chunks);
    hipLaunchKernelGGL(chunkScanKernel, dim3(static_cast<unsigned>(chunks)), dim3(kScanThreads), 0,
                       stream, hist, count, sums);
    hipLaunchKernelGGL(keyScatterKernel, dim3(nb), dim3(kSortBlock), 0, stream, keys, n,
                       nKeys, hist, perm);
  }
  FwdArgs fa;
  fa.in = b;
  fa.out = d;
  fa.sc = sc;
  fa.geo = c->geo;
  fa.slabs = c->slabs;
  fa.F = c->field;
  fa.o = c->opts;
  fa.pc = c->pc;
  fa.nSurfaces = c->nSurfaces;
  fa.nSlabs = c->nSlabs;
  fa.ckpt = ws.ckpt;
  fa.skipAfter = ws.ckpt ? c->skipAfter : 0;
  fa.pad = 0;
  fa.rv.scale = c->revScale;
  fa.rv.threshold = c->revThreshold;
  fa.rv.all = c->revAll ? 1 : 0;
  fa.rv.hasTarget = withRef ? 1 : 0;
  fa.rv.target = c->refLine;
  fa.rv.rec = ws.rev;
  fa.rv.reached = ws.revReached;
  int32_t* queue = ws.queue;
  hipLaunchKernelGGL(zeroKernel, dim3(1), dim3(64), 0, stream, queue, 1);
  const int64_t waves = std::min(blocks, c->residentWaves);
  if (timing) KF_HIP(hipEventRecord(c->ev[0], stream));
  hipLaunchKernelGGL(forwardKernel(c),
                     dim3(static_cast<unsigned>(waves)), dim3(kBlock), 0, stream, fa, queue,
                     c->innerSteps, c->waitFrac, c->quarantine, perm);
  if (timing) KF_HIP(hipEventRecord(c->ev[1], stream));
  const int32_t revOn = (c->revEnabled && ws.rev) ? 1 : 0;
  hipLaunchKernelGGL(c->mbf ? smoothKernel<true> : smoothKernel<false>,
                     dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, stream, b, d, sc, perm,
                     fa.rv, revOn);
  if (revOn) {
    // the reverse filter of the tracks that take one, in place of their
    // smoothing, before the extrapolation
    hipLaunchKernelGGL(zeroKernel, dim3(1), dim3(64), 0, stream, queue, 1);
    hipLaunchKernelGGL(forwardKernel(c, true),
                       dim3(static_cast<unsigned>(waves)), dim3(kBlock), 0, stream, fa, queue,
                       c->innerSteps, c->waitFrac, 0, perm);
  }
  if (withRef) {
    if (c->sympy && c->mixed)
      hipLaunchKernelGGL((referenceKernel<true, true>), dim3(static_cast<unsigned>(blocks)),
                         dim3(kBlock), 0, stream, d, sc, c->geo, c->field, c->extOpts,
                         c->refLine, c->refStrategy, n, in->seq_offsets, perm);
    else if (c->sympy)
      hipLaunchKernelGGL((referenceKernel<false, true>), dim3(static_cast<unsigned>(blocks)),
                         dim3(kBlock), 0, stream, d, sc, c->geo, c->field, c->extOpts,
                         c->refLine, c->refStrategy, n, in->seq_offsets, perm);
    else if (c->mixed)
      hipLaunchKernelGGL(referenceKernel<true>, dim3(static_cast<unsigned>(blocks)), dim3(kBlock),
                         0, stream, d, sc, c->geo, c->field, c->extOpts, c->refLine,
                         c->refStrategy, n, in->seq_offsets, perm);
    else
      hipLaunchKernelGGL(referenceKernel<false>, dim3(static_cast<unsigned>(blocks)),
                         dim3(kBlock), 0, stream, d, sc, c->geo, c->field, c->extOpts,
                         c->refLine, c->refStrategy, n, in->seq_offsets, perm);
  }
  const hipError_t launchErr = hipGetLastError();
  if (graph) {
    hipGraph_t graph = nullptr;
    const hipError_t endErr = hipStreamEndCapture(stream, &graph);
    if (launchErr != hipSuccess || endErr != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(ACTS_KF_ERR_DEVICE, std::string("graph capture failed: ") +
                                          hipGetErrorString(launchErr != hipSuccess ? launchErr
                                                                                    : endErr));
    }
    const hipError_t instErr = hipGraphInstantiate(&c->graphExec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (instErr != hipSuccess) {
      c->graphExec = nullptr;
      return fail(ACTS_KF_ERR_DEVICE,
                  std::string("hipGraphInstantiate: ") + hipGetErrorString(instErr));
    }
    c->graphKey = key;
    KF_HIP(hipGraphLaunch(c->graphExec, stream));
    return ACTS_KF_OK;
  }
  if (launchErr != hipSuccess)
    return fail(ACTS_KF_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(launchErr));
  if (timing) {
    KF_HIP(hipEventRecord(c->ev[2], stream));
    c->timed = true;
  }
  return ACTS_KF_OK;
}

int acts_kf_fit_batch_device(acts_kf_ctx* c, const acts_kf_batch_in* in,
                             const acts_kf_batch_out* out, void* stream) {
  if (!c || !in || !out) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "null argument");
  if (in->n_tracks < 0) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "negative n_tracks");
  if (in->n_tracks > 0 &&
      (!in->start_params || !in->start_cov || !in->start_plane || !in->seq_offsets ||
       !in->seq_surfaces || !in->meas_offsets || !in->meas_surfaces || !in->meas_values ||
       !in->meas_cov))
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "null input array");
  KF_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  int32_t maxLen = in->max_sequence_length;
  if (maxLen <= 0 && in->n_tracks > 0) {
    if (c->lenBuf.ensure(sizeof(int32_t))) return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "alloc");
    KF_HIP(hipMemsetAsync(c->lenBuf.p, 0, sizeof(int32_t), s));
    const int64_t nb = (in->n_tracks + 255) / 256;
    hipLaunchKernelGGL(maxLengthKernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s,
                       in->seq_offsets, in->n_tracks, static_cast<int32_t*>(c->lenBuf.p));
    KF_HIP(hipGetLastError());
    KF_HIP(hipMemcpyAsync(&maxLen, c->lenBuf.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    KF_HIP(hipStreamSynchronize(s));
  }
  if (in->n_tracks == 0) return ACTS_KF_OK;
  const WorkspaceSizes w = workspaceSizes(c, in->n_tracks, maxLen);
  Workspace ws;
  if (const int r = contextWorkspace(c, w, ws)) return r;
  if (const int r = launchFit(c, in, out, s, w, ws, c->useGraph)) return r;
  return markQueued(c, s);
}

int acts_kf_fit_batch_device_noalloc(acts_kf_ctx* c, const acts_kf_batch_in* in,
                                     const acts_kf_batch_out* out, void* stream) {
  if (!c || !in || !out) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "null argument");
  const int64_t n = in->n_tracks;
  if (n < 0) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "negative n_tracks");
  if (n == 0) return ACTS_KF_OK;
  if (!in->start_params || !in->start_cov || !in->start_plane || !in->seq_offsets ||
      !in->seq_surfaces || !in->meas_offsets || !in->meas_surfaces || !in->meas_values ||
      !in->meas_cov)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "null input array");
  if (n > kMaxBatch) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "batch too large");
  if (!c->staticWs)
    return fail(ACTS_KF_ERR_DEVICE, "acts_kf_fit_batch_device_noalloc: no static workspace");
  KF_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  // sections of the static workspace, 256-byte aligned: the sequence-length
  // slot, then the fit's workspace
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  char* base = static_cast<char*>(c->staticWs);
  int32_t maxLen = in->max_sequence_length;
  if (maxLen <= 0) {
    int32_t* len = reinterpret_cast<int32_t*>(base);
    hipLaunchKernelGGL(zeroKernel, dim3(1), dim3(64), 0, s, len, 1);
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(maxLengthKernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s,
                       in->seq_offsets, n, len);
    KF_HIP(hipGetLastError());
    KF_HIP(hipMemcpyAsync(&maxLen, len, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    KF_HIP(hipStreamSynchronize(s));
  }
  const WorkspaceSizes w = workspaceSizes(c, n, maxLen);
  const size_t sizes[] = {w.queue, w.track, w.perm, w.hist, w.ref, w.rev, w.i, w.d};
  size_t need = 256;
  for (size_t b : sizes) need += al(b);
  if (need > kStaticWorkspaceBytes)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT,
                "acts_kf_fit_batch_device_noalloc: the batch needs " + std::to_string(need) +
                    " bytes of workspace (748 per track and state slot), more than the " +
                    std::to_string(kStaticWorkspaceBytes) +
                    " of the static workspace; split it or use acts_kf_fit_batch_device");
  size_t off = 256;
  auto take = [&](size_t bytes) {
    void* p = bytes ? base + off : nullptr;
    off += al(bytes);
    return p;
  };
  Workspace ws;
  ws.queue = static_cast<int32_t*>(take(w.queue));
  ws.track = static_cast<int32_t*>(take(w.track));
  ws.perm = static_cast<int32_t*>(take(w.perm));
  ws.hist = static_cast<int32_t*>(take(w.hist));
  ws.ref = static_cast<double*>(take(w.ref));
  if (w.rev) reverseSections(static_cast<char*>(take(w.rev)), w, ws);
  ws.i = static_cast<int32_t*>(take(w.i));
  ws.d = static_cast<double*>(take(w.d));
  ws.ckpt = nullptr;  // no transport skipping: its checkpoints are context memory
  if (const int r = launchFit(c, in, out, s, w, ws, false)) return r;
  return markQueued(c, s);
}

int acts_kf_fit_batch(acts_kf_ctx* c, const acts_kf_batch_in* in,
                      const acts_kf_batch_out* out, void* stream) {
  if (!c || !in || !out) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "null argument");
  const int64_t n = in->n_tracks;
  if (n < 0) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "negative n_tracks");
  if (n == 0) return ACTS_KF_OK;
  if (!in->start_params || !in->start_cov || !in->start_plane || !in->seq_offsets ||
      !in->seq_surfaces || !in->meas_offsets || !in->meas_surfaces || !in->meas_values ||
      !in->meas_cov)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "null input array");
  if (!offsetsValid(in->seq_offsets, n) || !offsetsValid(in->meas_offsets, n))
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT,
                "acts_kf_fit_batch: seq_offsets and meas_offsets must start at 0 and not decrease");
  const int64_t nSeq = in->seq_offsets[n];
  const int64_t nMeas = in->meas_offsets[n];
  if (in->meas_dims || in->meas_indices) {
    for (int64_t m = 0; m < nMeas; ++m) {
      const int d = in->meas_dims ? in->meas_dims[m] : 2;
      const int a = in->meas_indices ? in->meas_indices[2 * m] : 0;
      const int b = in->meas_indices ? in->meas_indices[2 * m + 1] : 1;
      if ((d != 1 && d != 2) || a > 5 || (d == 2 && (b > 5 || b == a)))
        return fail(ACTS_KF_ERR_INVALID_ARGUMENT,
                    "acts_kf_fit_batch: measurement " + std::to_string(m) +
                        " has an unsupported subspace (dimension 1 or 2 on distinct "
                        "bound parameters 0..5)");
    }
  }
  KF_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  int32_t maxLen = 0;
  for (int64_t t = 0; t < n; ++t)
    maxLen = std::max<int32_t>(maxLen, static_cast<int32_t>(in->seq_offsets[t + 1] -
                                                            in->seq_offsets[t]));
  // input staging layout (8-byte aligned sections)
  struct Sec {
    const void* src;
    size_t bytes;
  };
  const Sec ins[] = {
      {in->start_params, sizeof(double) * 6 * n},  {in->start_cov, sizeof(double) * 36 * n},
      {in->start_plane, sizeof(double) * 12 * n},  {in->seq_offsets, sizeof(int64_t) * (n + 1)},
      {in->seq_surfaces, sizeof(int32_t) * nSeq},  {in->meas_offsets, sizeof(int64_t) * (n + 1)},
      {in->meas_surfaces, sizeof(int32_t) * nMeas}, {in->meas_values, sizeof(double) * 2 * nMeas},
      {in->meas_cov, sizeof(double) * 4 * nMeas},
      {in->meas_dims, in->meas_dims ? sizeof(uint8_t) * nMeas : 0},
      {in->meas_indices, in->meas_indices ? sizeof(uint8_t) * 2 * nMeas : 0}};
  constexpr int nIn = sizeof(ins) / sizeof(ins[0]);
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  size_t total = 0;
  for (const Sec& q : ins) total += al(q.bytes);
  if (c->inBuf.ensure(total)) return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "input staging alloc");
  char* base = static_cast<char*>(c->inBuf.p);
  void* dptr[nIn];
  size_t off = 0;
  for (int i = 0; i < nIn; ++i) {
    dptr[i] = base + off;
    if (ins[i].bytes)
      KF_HIP(hipMemcpyAsync(dptr[i], ins[i].src, ins[i].bytes, hipMemcpyHostToDevice, s));
    off += al(ins[i].bytes);
  }
  acts_kf_batch_in din = *in;
  din.start_params = static_cast<const double*>(dptr[0]);
  din.start_cov = static_cast<const double*>(dptr[1]);
  din.start_plane = static_cast<const double*>(dptr[2]);
  din.seq_offsets = static_cast<const int64_t*>(dptr[3]);
  din.seq_surfaces = static_cast<const int32_t*>(dptr[4]);
  din.meas_offsets = static_cast<const int64_t*>(dptr[5]);
  din.meas_surfaces = static_cast<const int32_t*>(dptr[6]);
  din.meas_values = static_cast<const double*>(dptr[7]);
  din.meas_cov = static_cast<const double*>(dptr[8]);
  din.meas_dims = in->meas_dims ? static_cast<const uint8_t*>(dptr[9]) : nullptr;
  din.meas_indices = in->meas_indices ? static_cast<const uint8_t*>(dptr[10]) : nullptr;
  din.max_sequence_length = maxLen;

  // output staging: the same arrays on the device, copied back afterwards
  struct OSec {
    void* host;
    size_t bytes;
  };
  const OSec outs[] = {
      {out->status, sizeof(int32_t) * n},          {out->n_states, sizeof(int32_t) * n},
      {out->chi2, sizeof(float) * n},              {out->ndf, sizeof(int32_t) * n},
      {out->n_measurements, sizeof(int32_t) * n},  {out->n_holes, sizeof(int32_t) * n},
      {out->n_outliers, sizeof(int32_t) * n},      {out->n_steps, sizeof(int32_t) * n},
      {out->n_step_trials, sizeof(int32_t) * n},   {out->state_surface, sizeof(int32_t) * nSeq},
      {out->state_flags, sizeof(uint32_t) * nSeq}, {out->state_path, sizeof(double) * nSeq},
      {out->state_chi2, sizeof(float) * nSeq},     {out->predicted, sizeof(double) * 6 * nSeq},
      {out->predicted_cov, sizeof(double) * 36 * nSeq},
      {out->filtered, sizeof(double) * 6 * nSeq},  {out->filtered_cov, sizeof(double) * 36 * nSeq},
      {out->smoothed, sizeof(double) * 6 * nSeq},  {out->smoothed_cov, sizeof(double) * 36 * nSeq},
      {out->jacobian, sizeof(double) * 36 * nSeq},
      {out->reference_params, sizeof(double) * 6 * n},
      {out->reference_cov, sizeof(double) * 36 * n},
      {out->state_measurement, sizeof(int32_t) * nSeq},
      {out->reversed, sizeof(int32_t) * n}};
  constexpr int nOut = sizeof(outs) / sizeof(outs[0]);
  size_t totalO = 0;
  for (const OSec& q : outs)
    if (q.host) totalO += al(q.bytes);
  if (c->outBuf.ensure(totalO)) return fail(ACTS_KF_ERR_OUT_OF_MEMORY, "output staging alloc");
  char* ob = static_cast<char*>(c->outBuf.p);
  void* optr[nOut];
  off = 0;
  for (int i = 0; i < nOut; ++i) {
    optr[i] = outs[i].host ? ob + off : nullptr;
    if (outs[i].host) off += al(outs[i].bytes);
  }
  acts_kf_batch_out dout;
  dout.status = static_cast<int32_t*>(optr[0]);
  dout.n_states = static_cast<int32_t*>(optr[1]);
  dout.chi2 = static_cast<float*>(optr[2]);
  dout.ndf = static_cast<int32_t*>(optr[3]);
  dout.n_measurements = static_cast<int32_t*>(optr[4]);
  dout.n_holes = static_cast<int32_t*>(optr[5]);
  dout.n_outliers = static_cast<int32_t*>(optr[6]);
  dout.n_steps = static_cast<int32_t*>(optr[7]);
  dout.n_step_trials = static_cast<int32_t*>(optr[8]);
  dout.state_surface = static_cast<int32_t*>(optr[9]);
  dout.state_flags = static_cast<uint32_t*>(optr[10]);
  dout.state_path = static_cast<double*>(optr[11]);
  dout.state_chi2 = static_cast<float*>(optr[12]);
  dout.predicted = static_cast<double*>(optr[13]);
  dout.predicted_cov = static_cast<double*>(optr[14]);
  dout.filtered = static_cast<double*>(optr[15]);
  dout.filtered_cov = static_cast<double*>(optr[16]);
  dout.smoothed = static_cast<double*>(optr[17]);
  dout.smoothed_cov = static_cast<double*>(optr[18]);
  dout.jacobian = static_cast<double*>(optr[19]);
  dout.reference_params = static_cast<double*>(optr[20]);
  dout.reference_cov = static_cast<double*>(optr[21]);
  dout.state_measurement = static_cast<int32_t*>(optr[22]);
  dout.reversed = static_cast<int32_t*>(optr[23]);
  const WorkspaceSizes w = workspaceSizes(c, n, maxLen);
  Workspace ws;
  if (const int r = contextWorkspace(c, w, ws)) return r;
  if (const int r = launchFit(c, &din, &dout, s, w, ws, c->useGraph)) return r;
  for (int i = 0; i < nOut; ++i)
    if (outs[i].host && outs[i].bytes)
      KF_HIP(hipMemcpyAsync(outs[i].host, optr[i], outs[i].bytes, hipMemcpyDeviceToHost, s));
  return markQueued(c, s);
}

int acts_kf_track_states_device(acts_kf_ctx* c, const acts_kf_batch_in* in,
                                const acts_kf_batch_out* f, const double* truth,
                                const acts_kf_track_states_out* out, void* stream) {
  if (!c || !in || !f || !out)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_track_states_device: null argument");
  if (in->n_tracks < 0 || out->n_slots < 0)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_track_states_device: negative size");
  if (in->n_tracks == 0 || out->n_slots == 0) return ACTS_KF_OK;
  if (!in->seq_offsets || !in->meas_offsets || !in->meas_values || !in->meas_cov)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT,
                "acts_kf_track_states_device: batch offsets or measurements are null");
  if (!f->status || !f->n_states || !f->state_surface || !f->state_flags || !f->state_path ||
      !f->state_chi2 || !f->state_measurement || !f->predicted || !f->predicted_cov ||
      !f->filtered || !f->filtered_cov || !f->smoothed || !f->smoothed_cov)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT,
                "acts_kf_track_states_device: a required fit output is null");
  if (!out->columns || !out->ints)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "acts_kf_track_states_device: output is null");
  KF_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  TsArgs a;
  a.nTracks = in->n_tracks;
  a.nSlots = out->n_slots;
  a.nSurfaces = c->nSurfaces;
  a.pad = 0;
  a.seqOff = in->seq_offsets;
  a.measOff = in->meas_offsets;
  a.measVal = in->meas_values;
  a.measCov = in->meas_cov;
  a.measDims = in->meas_dims;
  a.measIdx = in->meas_indices;
  a.status = f->status;
  a.nStates = f->n_states;
  a.stateSurf = f->state_surface;
  a.stateFlags = f->state_flags;
  a.statePath = f->state_path;
  a.stateChi2 = f->state_chi2;
  a.stateMeas = f->state_measurement;
  a.par[0] = f->predicted;
  a.par[1] = f->filtered;
  a.par[2] = f->smoothed;
  a.cov[0] = f->predicted_cov;
  a.cov[1] = f->filtered_cov;
  a.cov[2] = f->smoothed_cov;
  a.truth = truth;
  a.geo = c->geo;
  a.cols = out->columns;
  a.ints = out->ints;
  const int64_t blocks = (out->n_slots + kTsBlock - 1) / kTsBlock;
  hipLaunchKernelGGL((a.measDims || a.measIdx) ? trackStatesKernel<true> : trackStatesKernel<false>,
                     dim3(static_cast<unsigned>(blocks)), dim3(kTsBlock), 0, s, a);
  KF_HIP(hipGetLastError());
  return markQueued(c, s);
}

int acts_kf_kernel_times(acts_kf_ctx* c, double* forwardMs, double* smootherMs) {
  if (!c || !forwardMs || !smootherMs) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "null argument");
  if (!c->timing || !c->timed)
    return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "no timed fit (set options.kernel_timing)");
  KF_HIP(hipSetDevice(c->device));
  KF_HIP(hipEventSynchronize(c->ev[2]));
  float a = 0, b = 0;
  KF_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
  KF_HIP(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
  *forwardMs = a;
  *smootherMs = b;
  return ACTS_KF_OK;
}

int acts_kf_sync(acts_kf_ctx* c) {
  if (!c) return fail(ACTS_KF_ERR_INVALID_ARGUMENT, "null context");
  if (!c->pending) return ACTS_KF_OK;
  KF_HIP(hipSetDevice(c->device));
  KF_HIP(hipEventSynchronize(c->done));
  c->pending = false;
  return ACTS_KF_OK;
}

void acts_kf_destroy(acts_kf_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->scratchD.release();
  c->scratchI.release();
  c->scratchT.release();
  c->refBuf.release();
  c->ckptBuf.release();
  c->revBuf.release();
  c->lenBuf.release();
  c->permBuf.release();
  c->histBuf.release();
  c->queueBuf.release();
  c->inBuf.release();
  c->outBuf.release();
  if (c->geo) (void)hipFree(c->geo);
  if (c->slabs) (void)hipFree(c->slabs);
  if (c->fieldValues) (void)hipFree(c->fieldValues);
  if (c->fieldValuesF) (void)hipFree(c->fieldValuesF);
  if (c->fieldCells) (void)hipFree(c->fieldCells);
  if (c->fieldCellsF) (void)hipFree(c->fieldCellsF);
  if (c->graphExec) (void)hipGraphExecDestroy(c->graphExec);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->pending) (void)hipEventSynchronize(c->done);
  if (c->done) (void)hipEventDestroy(c->done);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
