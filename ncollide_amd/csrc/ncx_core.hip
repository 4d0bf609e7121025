// ncx_core.hip -- context, errors, device-wide scan, LSD radix sort and the
// Morton-sorted LBVH (Karras 2012 hierarchy, atomic bottom-up refit) shared by
// the broad phase and TriMesh ray casting.
#include "ncx_internal.h"
#include <cstdio>
#include <mutex>
#include <algorithm>

static thread_local std::string g_last_error;
void ncx_set_error(const std::string& msg) { g_last_error = msg; }
extern "C" const char* ncx_last_error(void) { return g_last_error.c_str(); }

extern "C" int ncx_init(ncx_ctx** out, int device, void* stream) {
  NCX_REQUIRE(out != nullptr, "ncx_init: null output");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    ncx_set_error("ncx_init: no HIP device visible");
    return NCX_E_NODEVICE;
  }
  NCX_REQUIRE(device >= 0 && device < count, "ncx_init: device index out of range");
  NCX_HIP_CHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  NCX_HIP_CHECK(hipGetDeviceProperties(&prop, device));
  int rc = rc_scratch_setup(device);
  if (rc) return rc;
  ncx_ctx* c = new ncx_ctx;
  c->device = device;
  c->stream = (hipStream_t)stream;
  c->num_cus = prop.multiProcessorCount;
  *out = c;
  return NCX_OK;
}

extern "C" int ncx_free(ncx_ctx* ctx) {
  delete ctx;
  return NCX_OK;
}

extern "C" int ncx_set_stream(ncx_ctx* ctx, void* stream) {
  NCX_REQUIRE(ctx, "ncx_set_stream: null context");
  ctx->stream = (hipStream_t)stream;
  return NCX_OK;
}

extern "C" int ncx_synchronize(ncx_ctx* ctx) {
  NCX_REQUIRE(ctx, "ncx_synchronize: null context");
  NCX_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return NCX_OK;
}

// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// ------------------------------------------------------------------ exclusive scan
#define SCAN_BLOCK 256
#define SCAN_ITEMS 8
#define SCAN_TILE (SCAN_BLOCK * SCAN_ITEMS)

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                            size_t n, uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t wsum[SCAN_BLOCK / 64];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    size_t e = base + i;
    v[i] = e < n ? in[e] : 0u;
    tot += v[i];
  }
  uint32_t incl = wave_incl_scan(tot);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  uint32_t wpre = 0, bsum = 0;
#pragma unroll
  for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
    uint32_t s = wsum[w];
    if (w < wid) wpre += s;
    bsum += s;
  }
  uint32_t run = wpre + incl - tot;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    size_t e = base + i;
    if (e < n) out[e] = run;
    run += v[i];
  }
  if (threadIdx.x == 0 && block_sums) block_sums[blockIdx.x] = bsum;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(uint32_t* __restrict__ out, size_t n,
                                                          const uint32_t* __restrict__ block_offs) {
  const uint32_t add = block_offs[blockIdx.x];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_BLOCK) {
    size_t e = base + i;
    if (e < n) out[e] += add;
  }
}

__global__ void k_total(const uint32_t* in_last, const uint32_t* out_last, uint32_t* total) {
  *total = *in_last + *out_last;
}

static int scan_rec(ncx_ctx* ctx, const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp, size_t tmp_words) {
  unsigned nb = ceil_div(n, SCAN_TILE);
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_BLOCK), 0, ctx->stream, in, out, n, (uint32_t*)nullptr);
    NCX_HIP_CHECK(hipGetLastError());
    return NCX_OK;
  }
  NCX_REQUIRE(tmp_words >= (size_t)nb * 2, "scan: scratch too small");
  uint32_t* sums = tmp;
  uint32_t* scanned = tmp + nb;
  uint32_t* rest = tmp + 2 * (size_t)nb;
  hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(SCAN_BLOCK), 0, ctx->stream, in, out, n, sums);
  NCX_HIP_CHECK(hipGetLastError());
  int rc = scan_rec(ctx, sums, scanned, nb, rest, tmp_words - 2 * (size_t)nb);
  if (rc) return rc;
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(SCAN_BLOCK), 0, ctx->stream, out, n, scanned);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

int scan_exclusive_u32(ncx_ctx* ctx, const uint32_t* in, uint32_t* out, size_t n, DevBuf& tmp, uint32_t* total_dev) {
  if (n == 0) return NCX_OK;
  NCX_REQUIRE(in != out, "scan_exclusive_u32: in-place scan not supported");
  size_t words = 0, m = n;
  while (m > SCAN_TILE) { m = ceil_div(m, SCAN_TILE); words += 2 * m; }
  words += 2 * SCAN_TILE + 64;
  int rc = tmp.reserve(words * 4);
  if (rc) return rc;
  rc = scan_rec(ctx, in, out, n, tmp.as<uint32_t>(), words);
  if (rc) return rc;
  if (total_dev) {
    hipLaunchKernelGGL(k_total, dim3(1), dim3(1), 0, ctx->stream, in + n - 1, out + n - 1, total_dev);
    NCX_HIP_CHECK(hipGetLastError());
  }
  return NCX_OK;
}

// ------------------------------------------------------------------ LSD radix sort (8-bit digits)
#define RS_BLOCK 256
#define RS_ITEMS 8
#define RS_TILE (RS_BLOCK * RS_ITEMS)

__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint32_t* __restrict__ keys, size_t n, int shift,
                                                       uint32_t nbits, uint32_t* __restrict__ hist, unsigned nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t mask = (1u << nbits) - 1u;
  const size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    size_t e = base + (size_t)r * RS_BLOCK + threadIdx.x;
    if (e < n) atomicAdd(&h[(keys[e] >> shift) & mask], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const uint32_t* __restrict__ keys_in,
                                                          const uint32_t* __restrict__ vals_in,
                                                          uint32_t* __restrict__ keys_out,
                                                          uint32_t* __restrict__ vals_out, size_t n, int shift,
                                                          uint32_t nbits, const uint32_t* __restrict__ offs,
                                                          unsigned nblocks) {
  __shared__ uint32_t base_d[256];
  __shared__ uint32_t run_d[256];
  __shared__ uint32_t cnt[RS_BLOCK / 64][256];
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  base_d[tid] = offs[(size_t)tid * nblocks + blockIdx.x];
  run_d[tid] = 0;
  const uint32_t mask = (1u << nbits) - 1u;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const size_t tbase = (size_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
#pragma unroll
    for (int w = 0; w < RS_BLOCK / 64; ++w) cnt[w][tid] = 0;
    __syncthreads();
    size_t e = tbase + (size_t)r * RS_BLOCK + tid;
    bool valid = e < n;
    uint32_t key = valid ? keys_in[e] : 0u;
    uint32_t val = (valid && vals_in) ? vals_in[e] : 0u;
    uint32_t d = (key >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (uint32_t b = 0; b < nbits; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t rank = __popcll(peers & lt);
    if (valid && rank == 0) cnt[wid][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pre = run_d[d];
      for (int w = 0; w < wid; ++w) pre += cnt[w][d];
      uint32_t dst = base_d[d] + pre + rank;
      keys_out[dst] = key;
      if (vals_out) vals_out[dst] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < RS_BLOCK / 64; ++w) add += cnt[w][tid];
    run_d[tid] += add;
    __syncthreads();
  }
}

int radix_sort_u32(ncx_ctx* ctx, uint32_t* keys, uint32_t* vals, size_t n, int begin_bit, int end_bit,
                   DevBuf& tmp_keys, DevBuf& tmp_vals, DevBuf& hist, DevBuf& scan_tmp) {
  if (n <= 1 || end_bit <= begin_bit) return NCX_OK;
  NCX_REQUIRE(n < 0xFFFFFFFFull, "radix_sort_u32: too many keys");
  unsigned nb = ceil_div(n, RS_TILE);
  int rc;
  if ((rc = tmp_keys.reserve(n * 4))) return rc;
  if (vals && (rc = tmp_vals.reserve(n * 4))) return rc;
  if ((rc = hist.reserve((size_t)256 * nb * 4 * 2))) return rc;
  uint32_t* h = hist.as<uint32_t>();
  uint32_t* hs = h + (size_t)256 * nb;
  uint32_t *ki = keys, *vi = vals, *ko = tmp_keys.as<uint32_t>(), *vo = vals ? tmp_vals.as<uint32_t>() : nullptr;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    uint32_t nbits = (uint32_t)((end_bit - shift) < 8 ? (end_bit - shift) : 8);
    hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_BLOCK), 0, ctx->stream, ki, n, shift, nbits, h, nb);
    NCX_HIP_CHECK(hipGetLastError());
    if ((rc = scan_exclusive_u32(ctx, h, hs, (size_t)256 * nb, scan_tmp, nullptr))) return rc;
    hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_BLOCK), 0, ctx->stream, ki, vi, ko, vo, n, shift, nbits, hs, nb);
    NCX_HIP_CHECK(hipGetLastError());
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  if (ki != keys) {
    NCX_HIP_CHECK(hipMemcpyAsync(keys, ki, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    if (vals) NCX_HIP_CHECK(hipMemcpyAsync(vals, vi, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return NCX_OK;
}

int radix_sort_pairs64(ncx_ctx* ctx, uint32_t* hi, uint32_t* lo, size_t n, int hi_bits, int lo_bits, DevBuf& t1,
                       DevBuf& t2, DevBuf& t3, DevBuf& hist, DevBuf& scan_tmp) {
  (void)t3;
  int rc = radix_sort_u32(ctx, lo, hi, n, 0, lo_bits, t1, t2, hist, scan_tmp);
  if (rc) return rc;
  return radix_sort_u32(ctx, hi, lo, n, 0, hi_bits, t1, t2, hist, scan_tmp);
}

// ------------------------------------------------------------------ LBVH
__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__global__ void k_bounds_init(uint32_t* b) {
  int t = threadIdx.x;
  if (t < 3) b[t] = 0xFFFFFFFFu;
  else if (t < 6) b[t] = 0u;
}

__global__ __launch_bounds__(256) void k_centroid_bounds(const float* __restrict__ minx, const float* __restrict__ miny,
                                                          const float* __restrict__ minz, const float* __restrict__ maxx,
                                                          const float* __restrict__ maxy, const float* __restrict__ maxz,
                                                          size_t n, uint32_t* __restrict__ b) {
  float lo[3] = {NCX_FLT_MAX, NCX_FLT_MAX, NCX_FLT_MAX}, hi[3] = {-NCX_FLT_MAX, -NCX_FLT_MAX, -NCX_FLT_MAX};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float c[3] = {(minx[i] + maxx[i]) * 0.5f, (miny[i] + maxy[i]) * 0.5f, (minz[i] + maxz[i]) * 0.5f};
#pragma unroll
    for (int d = 0; d < 3; ++d) { lo[d] = fminf(lo[d], c[d]); hi[d] = fmaxf(hi[d], c[d]); }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    for (int off = 32; off > 0; off >>= 1) {
      lo[d] = fminf(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], off, 64));
    }
  }
  __shared__ float red[6][4];
  const int w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) { red[d][w] = lo[d]; red[3 + d][w] = hi[d]; }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    int d = threadIdx.x;
    float v = red[d][0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) v = d < 3 ? fminf(v, red[d][q]) : fmaxf(v, red[d][q]);
    if (d < 3) atomicMin(&b[d], f2ord(v));
    else atomicMax(&b[d], f2ord(v));
  }
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ minx, const float* __restrict__ miny,
                                                 const float* __restrict__ minz, const float* __restrict__ maxx,
                                                 const float* __restrict__ maxy, const float* __restrict__ maxz,
                                                 size_t n, const uint32_t* __restrict__ b, uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float c[3] = {(minx[i] + maxx[i]) * 0.5f, (miny[i] + maxy[i]) * 0.5f, (minz[i] + maxz[i]) * 0.5f};
  // One scale for all three axes (the largest centroid extent): a flat set of primitives (a
  // heightfield) then splits along its long axes first instead of cutting it into thin layers
  // that each span the whole set.
  float ext = 0.f;
#pragma unroll
  for (int d = 0; d < 3; ++d) ext = fmaxf(ext, ord2f(b[3 + d]) - ord2f(b[d]));
  uint32_t q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float lo = ord2f(b[d]);
    float u = ext > 0.f ? (c[d] - lo) / ext : 0.f;
    float s = fminf(fmaxf(u * 1024.f, 0.f), 1023.f);
    q[d] = (uint32_t)s;
  }
  keys[i] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
  vals[i] = (uint32_t)i;
}

__device__ __forceinline__ int lbvh_delta(const uint32_t* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = keys[i], b = keys[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees and k-d trees".
__global__ __launch_bounds__(256) void k_hierarchy(const uint32_t* __restrict__ keys, int n, LNode* __restrict__ nodes,
                                                    uint32_t* __restrict__ parent, uint32_t* __restrict__ first_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (lbvh_delta(keys, n, i, i + 1) - lbvh_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = lbvh_delta(keys, n, i, i - d);
  int lmax = 2;
  while (lbvh_delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (lbvh_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = lbvh_delta(keys, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (lbvh_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + (d < 0 ? d : 0);
  int first = i < j ? i : j, last = i < j ? j : i;
  uint32_t cl, cr;
  if (first == gamma) { cl = (uint32_t)gamma | NCX_LEAF_BIT; parent[(n - 1) + gamma] = (uint32_t)i; }
  else { cl = (uint32_t)gamma; parent[gamma] = (uint32_t)i; }
  if (last == gamma + 1) { cr = (uint32_t)(gamma + 1) | NCX_LEAF_BIT; parent[(n - 1) + gamma + 1] = (uint32_t)i | NCX_LEAF_BIT; }
  else { cr = (uint32_t)(gamma + 1); parent[gamma + 1] = (uint32_t)i | NCX_LEAF_BIT; }
  // NCX_LEAF_BIT on a parent entry marks "right child of".
  nodes[i].child_l = cl;
  nodes[i].child_r = cr;
  nodes[i].last_l = (uint32_t)gamma;
  nodes[i].last_r = (uint32_t)last;
  first_out[i] = (uint32_t)first;
}

__global__ void k_zero_u32(uint32_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int fill_u32(ncx_ctx* ctx, uint32_t* p, size_t n, uint32_t v) {
  if (n == 0) return NCX_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, p, n, v);
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

// ------------------------------------------------------------------ refit
// Node boxes live in the parent's child slots, so refitting writes every node's box into its
// parent.  Stage 1 (one workgroup per REFIT_B consecutive leaves) refits every node whose leaf
// range lies inside the workgroup's leaves, with flags and boxes in LDS (the second child to
// arrive merges).  Stage 2 (one workgroup) refits the few nodes whose range spans workgroups,
// deepest level first, with a barrier between levels.  No device-scope synchronisation is needed.
#define REFIT_B 1024

__device__ __forceinline__ void expand_box(float expand, float b[6]) {
  if (expand > 0.f) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b[d] -= expand * (1.f + fabsf(b[d]));
      b[3 + d] += expand * (1.f + fabsf(b[3 + d]));
    }
  }
}

__device__ __forceinline__ void leaf_box(const float* const* box, uint32_t p, float expand, float b[6]) {
#pragma unroll
  for (int d = 0; d < 6; ++d) b[d] = box[d][p];
  expand_box(expand, b);
}

__device__ __forceinline__ void put_slot(LNode* nd, bool right, const float b[6]) {
  float* f = (float*)nd + (right ? 6 : 0);
#pragma unroll
  for (int k = 0; k < 6; ++k) f[k] = b[k];
}

struct BoxPtrs {
  const float* p[6];
};

template <bool AOS>
__global__ __launch_bounds__(REFIT_B) void k_refit_local(BoxPtrs bx, const float4* __restrict__ aos,
                                                          const uint32_t* __restrict__ order, int n,
                                                          LNode* __restrict__ nodes,
                                                          const uint32_t* __restrict__ parent, float expand) {
  __shared__ uint32_t lflag[REFIT_B];
  __shared__ float lbox[REFIT_B][12];
  const int tid = threadIdx.x;
  lflag[tid] = 0u;
  __syncthreads();
  const int b0 = blockIdx.x * REFIT_B;
  const int k = b0 + tid;
  if (k >= n) return;
  float b[6];
  if constexpr (AOS) {
    const uint32_t p = order[k];
    const float4 a0 = aos[2 * (size_t)p], a1 = aos[2 * (size_t)p + 1];
    b[0] = a0.x; b[1] = a0.y; b[2] = a0.z; b[3] = a0.w; b[4] = a1.x; b[5] = a1.y;
    expand_box(expand, b);
  } else {
    leaf_box(bx.p, order[k], expand, b);
  }
  uint32_t pe = parent[(n - 1) + k];
  for (;;) {
    uint32_t par = pe & NCX_PARENT_MASK;
    bool right = (pe & NCX_RIGHT_BIT) != 0;
    put_slot(&nodes[par], right, b);
    // The span bit (the parent's leaf range leaves this workgroup's) saves loading the range.
    if (pe & NCX_SPAN_BIT) return;  // stage 2
    int slot = (int)par - b0;
#pragma unroll
    for (int q = 0; q < 6; ++q) lbox[slot][(right ? 6 : 0) + q] = b[q];
    uint32_t old = __hip_atomic_fetch_add(&lflag[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == 0u) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b[d] = fminf(lbox[slot][d], lbox[slot][6 + d]);
      b[3 + d] = fmaxf(lbox[slot][3 + d], lbox[slot][9 + d]);
    }
    if (par == 0) return;
    pe = parent[par];
  }
}

// Stage-2 entries held in LDS by k_refit_upper (64 KiB); entries past it are read from memory.
#define UPPER_LDS 3072
#define UPPER_THREADS 1024
#define UPPER_NONE 0xFFFFFFFFu

__device__ __forceinline__ void merge_slots(const float* f, float b[6]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b[d] = fminf(f[d], f[6 + d]);
    b[3 + d] = fmaxf(f[3 + d], f[9 + d]);
  }
}

// Stage 2, one workgroup, level by level, deepest first.  Entry = (node, the node's parent entry,
// the stage-2 entry of its left child, of its right child; UPPER_NONE for a child refitted by
// stage 1).  Up to UPPER_LDS entries: every node's child slots are fetched at the start (one
// round trip for all levels), a stage-2 child's box comes from LDS, and a level costs a barrier;
// at the end every box is stored into its parent's slot for the traversals.  Past UPPER_LDS entries:
// the same levels through device memory, one round trip per level.
__global__ __launch_bounds__(UPPER_THREADS) void k_refit_upper(LNode* __restrict__ nodes, const uint4* __restrict__ upper,
                                                                const uint32_t* __restrict__ level_off) {
  __shared__ uint32_t off[LBVH_MAX_DEPTH + 1];
  __shared__ uint4 ent[UPPER_LDS];
  __shared__ float mb[UPPER_LDS][6];
  constexpr int J = UPPER_LDS / UPPER_THREADS;
  const uint32_t tid = threadIdx.x;
  if (tid <= LBVH_MAX_DEPTH) off[tid] = level_off[tid];
  __syncthreads();
  const uint32_t total = off[LBVH_MAX_DEPTH];
  if (total <= UPPER_LDS) {
    float slot[J][12];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const uint32_t i = tid + j * UPPER_THREADS;
      if (i < total) {
        const uint4 en = upper[i];
        ent[i] = en;
        const float4* f = (const float4*)&nodes[en.x];
        const float4 f0 = f[0], f1 = f[1], f2 = f[2];
        slot[j][0] = f0.x; slot[j][1] = f0.y; slot[j][2] = f0.z; slot[j][3] = f0.w;
        slot[j][4] = f1.x; slot[j][5] = f1.y; slot[j][6] = f1.z; slot[j][7] = f1.w;
        slot[j][8] = f2.x; slot[j][9] = f2.y; slot[j][10] = f2.z; slot[j][11] = f2.w;
      }
    }
    __syncthreads();
    for (int lv = 0; lv < (int)LBVH_MAX_DEPTH; ++lv) {
      const uint32_t a = off[lv], e = off[lv + 1];
      if (a == e) continue;  // uniform across the workgroup
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const uint32_t i = tid + j * UPPER_THREADS;
        if (i < a || i >= e) continue;
        const uint4 en = ent[i];
        if (en.z != UPPER_NONE)
          for (int q = 0; q < 6; ++q) slot[j][q] = mb[en.z][q];
        if (en.w != UPPER_NONE)
          for (int q = 0; q < 6; ++q) slot[j][6 + q] = mb[en.w][q];
        float b[6];
        merge_slots(slot[j], b);
        for (int q = 0; q < 6; ++q) mb[i][q] = b[q];
      }
      __syncthreads();
    }
    // the boxes go to their parents' slots once all levels are done: a store pending at a
    // level's barrier would make the barrier wait for it
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const uint32_t i = tid + j * UPPER_THREADS;
      if (i < total && ent[i].x != 0) {
        float b[6];
        for (int q = 0; q < 6; ++q) b[q] = mb[i][q];
        put_slot(&nodes[ent[i].y & NCX_PARENT_MASK], (ent[i].y & NCX_RIGHT_BIT) != 0, b);
      }
    }
    return;
  }
  for (int lv = 0; lv < (int)LBVH_MAX_DEPTH; ++lv) {
    const uint32_t a = off[lv], e = off[lv + 1];
    if (a == e) continue;
    for (uint32_t i = a + tid; i < e; i += UPPER_THREADS) {
      const uint4 en = upper[i];
      float b[6];
      merge_slots((const float*)&nodes[en.x], b);
      if (en.x != 0) put_slot(&nodes[en.y & NCX_PARENT_MASK], (en.y & NCX_RIGHT_BIT) != 0, b);
    }
    __syncthreads();
  }
}

// Stage-2 nodes (leaf range spanning refit workgroups): flag and depth (LBVH_MAX_DEPTH for the
// others), and the span bit in their children's parent entries.  No atomics: a few thousand
// nodes counting into 64 shared counters serialise.  Each parent entry is rewritten only by its
// parent's thread; the depth walks read entries concurrently and mask the bit off.
__global__ __launch_bounds__(256) void k_upper_mark(const LNode* __restrict__ nodes, uint32_t* __restrict__ parent,
                                                     const uint32_t* __restrict__ first, int n,
                                                     uint32_t* __restrict__ depth, uint32_t* __restrict__ flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const bool upper = first[i] / REFIT_B != nodes[i].last_r / REFIT_B;
  uint32_t d = LBVH_MAX_DEPTH;
  if (upper) {
    const uint32_t cl = nodes[i].child_l, cr = nodes[i].child_r;
    const uint32_t el = (cl & NCX_LEAF_BIT) ? (uint32_t)(n - 1) + (cl & ~NCX_LEAF_BIT) : cl;
    const uint32_t er = (cr & NCX_LEAF_BIT) ? (uint32_t)(n - 1) + (cr & ~NCX_LEAF_BIT) : cr;
    parent[el] |= NCX_SPAN_BIT;
    parent[er] |= NCX_SPAN_BIT;
    d = 0;
    uint32_t p = (uint32_t)i;
    while (p != 0 && d < LBVH_MAX_DEPTH - 1) {
      p = parent[p] & NCX_PARENT_MASK;
      d++;
    }
  }
  depth[i] = d;
  flag[i] = upper ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_upper_compact(const uint32_t* __restrict__ flag,
                                                        const uint32_t* __restrict__ pos,
                                                        const uint32_t* __restrict__ depth,
                                                        const uint32_t* __restrict__ parent, int m,
                                                        uint4* __restrict__ list) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && flag[i]) list[pos[i]] = make_uint4((uint32_t)i, depth[i], i != 0 ? parent[i] : 0u, 0u);
}

// One workgroup: counting sort of the compacted stage-2 nodes (node, depth, parent entry) by
// depth, deepest first, into k_refit_upper's entries, with level offsets off[lv] for lv =
// 0..LBVH_MAX_DEPTH (level lv holds depth LBVH_MAX_DEPTH - 1 - lv).  Order within a level is
// free: a level's nodes write distinct parent slots.  Each thread loads its entries of a chunk
// at once, then counts or places them, and records pos[] (node -> entry) for k_upper_link.
#define SORT_PER_THREAD 8
__global__ __launch_bounds__(1024) void k_upper_sort(const uint4* __restrict__ list, const uint32_t* __restrict__ total_p,
                                                      uint32_t* __restrict__ pos, uint4* __restrict__ upper,
                                                      uint32_t* __restrict__ off) {
  __shared__ uint32_t cnt[LBVH_MAX_DEPTH], cur[LBVH_MAX_DEPTH];
  const uint32_t tid = threadIdx.x;
  const uint32_t chunk = SORT_PER_THREAD * blockDim.x;
  if (tid < LBVH_MAX_DEPTH) cnt[tid] = 0u;
  __syncthreads();
  const uint32_t total = *total_p;
  for (uint32_t c0 = 0; c0 < total; c0 += chunk) {
    uint32_t d[SORT_PER_THREAD];
#pragma unroll
    for (int j = 0; j < SORT_PER_THREAD; ++j) {
      const uint32_t i = c0 + tid + j * blockDim.x;
      d[j] = i < total ? list[i].y : LBVH_MAX_DEPTH;
    }
#pragma unroll
    for (int j = 0; j < SORT_PER_THREAD; ++j)
      if (d[j] < LBVH_MAX_DEPTH) atomicAdd(&cnt[d[j]], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t o = 0;
    for (uint32_t lv = 0; lv < LBVH_MAX_DEPTH; ++lv) {
      const uint32_t dd = LBVH_MAX_DEPTH - 1 - lv;
      off[lv] = o;
      cur[dd] = o;
      o += cnt[dd];
    }
    off[LBVH_MAX_DEPTH] = o;
  }
  __syncthreads();
  for (uint32_t c0 = 0; c0 < total; c0 += chunk) {
    uint4 e[SORT_PER_THREAD];
#pragma unroll
    for (int j = 0; j < SORT_PER_THREAD; ++j) {
      const uint32_t i = c0 + tid + j * blockDim.x;
      e[j] = i < total ? list[i] : make_uint4(0u, LBVH_MAX_DEPTH, 0u, 0u);
    }
#pragma unroll
    for (int j = 0; j < SORT_PER_THREAD; ++j)
      if (e[j].y < LBVH_MAX_DEPTH) {
        const uint32_t slot = atomicAdd(&cur[e[j].y], 1u);
        upper[slot] = make_uint4(e[j].x, e[j].z, UPPER_NONE, UPPER_NONE);
        pos[e[j].x] = slot;
      }
  }
}

// Links each stage-2 entry to its children's entries (pos[] maps node -> entry); a grid of its
// own, so the dependent loads of all entries are in flight at once.
#define UPPER_LINK_BLOCKS 64
__global__ __launch_bounds__(256) void k_upper_link(const uint32_t* __restrict__ total_p, const LNode* __restrict__ nodes,
                                                     const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                                     uint4* __restrict__ upper) {
  const uint32_t total = *total_p;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < total; s += gridDim.x * blockDim.x) {
    const uint32_t p = upper[s].x;
    const uint32_t cl = nodes[p].child_l, cr = nodes[p].child_r;
    const uint32_t el = (!(cl & NCX_LEAF_BIT) && flag[cl]) ? pos[cl] : UPPER_NONE;
    const uint32_t er = (!(cr & NCX_LEAF_BIT) && flag[cr]) ? pos[cr] : UPPER_NONE;
    upper[s].z = el;
    upper[s].w = er;
  }
}

// Stage-2 node list for a new topology, on the device (no host synchronisation).
static int refit_prepare(ncx_ctx* ctx, LBVH& bvh) {
  const size_t m = bvh.n - 1;
  int rc;
  if ((rc = bvh.node_depth.reserve(m * 4)) || (rc = bvh.upper.reserve(m * 16)) ||
      (rc = bvh.upper_off.reserve(((size_t)LBVH_MAX_DEPTH + 2) * 4)) || (rc = bvh.flags.reserve(m * 4)) ||
      (rc = bvh.tmp_keys.reserve(m * 4)) || (rc = bvh.tmp_vals.reserve(m * 16)))
    return rc;
  uint32_t* total = bvh.upper_off.as<uint32_t>() + LBVH_MAX_DEPTH + 1;
  hipLaunchKernelGGL(k_upper_mark, dim3(ceil_div(m, 256)), dim3(256), 0, ctx->stream, bvh.nodes.as<LNode>(),
                     bvh.parent.as<uint32_t>(), bvh.first.as<uint32_t>(), (int)bvh.n, bvh.node_depth.as<uint32_t>(),
                     bvh.flags.as<uint32_t>());
  NCX_HIP_CHECK(hipGetLastError());
  if ((rc = scan_exclusive_u32(ctx, bvh.flags.as<uint32_t>(), bvh.tmp_keys.as<uint32_t>(), m, bvh.scan_tmp, total)))
    return rc;
  hipLaunchKernelGGL(k_upper_compact, dim3(ceil_div(m, 256)), dim3(256), 0, ctx->stream, bvh.flags.as<uint32_t>(),
                     bvh.tmp_keys.as<uint32_t>(), bvh.node_depth.as<uint32_t>(), bvh.parent.as<uint32_t>(), (int)m,
                     bvh.tmp_vals.as<uint4>());
  NCX_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_upper_sort, dim3(1), dim3(1024), 0, ctx->stream, bvh.tmp_vals.as<uint4>(), total,
                     bvh.tmp_keys.as<uint32_t>(), bvh.upper.as<uint4>(), bvh.upper_off.as<uint32_t>());
  NCX_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_upper_link, dim3(UPPER_LINK_BLOCKS), dim3(256), 0, ctx->stream, total, bvh.nodes.as<LNode>(),
                     bvh.flags.as<uint32_t>(), bvh.tmp_keys.as<uint32_t>(), bvh.upper.as<uint4>());
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

static int refit_launch(ncx_ctx* ctx, LBVH& bvh, const float* const box[6], float expand, const float4* aos) {
  size_t n = bvh.n;
  BoxPtrs bp;
  for (int d = 0; d < 6; ++d) bp.p[d] = box[d];
  if (aos)
    hipLaunchKernelGGL(k_refit_local<true>, dim3(ceil_div(n, REFIT_B)), dim3(REFIT_B), 0, ctx->stream, bp, aos,
                       bvh.order.as<uint32_t>(), (int)n, bvh.nodes.as<LNode>(), bvh.parent.as<uint32_t>(), expand);
  else
    hipLaunchKernelGGL(k_refit_local<false>, dim3(ceil_div(n, REFIT_B)), dim3(REFIT_B), 0, ctx->stream, bp, aos,
                       bvh.order.as<uint32_t>(), (int)n, bvh.nodes.as<LNode>(), bvh.parent.as<uint32_t>(), expand);
  NCX_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_refit_upper, dim3(1), dim3(UPPER_THREADS), 0, ctx->stream, bvh.nodes.as<LNode>(),
                     bvh.upper.as<uint4>(), bvh.upper_off.as<uint32_t>());
  NCX_HIP_CHECK(hipGetLastError());
  return NCX_OK;
}

int lbvh_build(ncx_ctx* ctx, LBVH& bvh, const float* const box[6], size_t n, float expand, const float4* aos) {
  int rc;
  NCX_REQUIRE(n <= (size_t)NCX_PARENT_MASK, "lbvh_build: too many primitives");
  bvh.n = n;
  if (n == 0) return NCX_OK;
  if ((rc = bvh.order.reserve(n * 4))) return rc;
  if ((rc = bvh.keys.reserve(n * 4))) return rc;
  if ((rc = bvh.bounds.reserve(64))) return rc;
  if ((rc = bvh.nodes.reserve((n > 1 ? n - 1 : 1) * sizeof(LNode)))) return rc;
  if ((rc = bvh.parent.reserve((2 * n) * 4))) return rc;
  if ((rc = bvh.flags.reserve(n * 4))) return rc;
  uint32_t* b = bvh.bounds.as<uint32_t>();
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, ctx->stream, b);
  unsigned gb = ceil_div(n, 256);
  if (gb > 256) gb = 256;
  hipLaunchKernelGGL(k_centroid_bounds, dim3(gb), dim3(256), 0, ctx->stream, box[0], box[1], box[2], box[3], box[4],
                     box[5], n, b);
  hipLaunchKernelGGL(k_morton, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, box[0], box[1], box[2], box[3],
                     box[4], box[5], n, b, bvh.keys.as<uint32_t>(), bvh.order.as<uint32_t>());
  NCX_HIP_CHECK(hipGetLastError());
  if ((rc = radix_sort_u32(ctx, bvh.keys.as<uint32_t>(), bvh.order.as<uint32_t>(), n, 0, 30, bvh.tmp_keys,
                           bvh.tmp_vals, bvh.hist, bvh.scan_tmp)))
    return rc;
  if (n == 1) return NCX_OK;
  if ((rc = bvh.first.reserve((n - 1) * 4))) return rc;
  hipLaunchKernelGGL(k_hierarchy, dim3(ceil_div(n - 1, 256)), dim3(256), 0, ctx->stream, bvh.keys.as<uint32_t>(),
                     (int)n, bvh.nodes.as<LNode>(), bvh.parent.as<uint32_t>(), bvh.first.as<uint32_t>());
  NCX_HIP_CHECK(hipGetLastError());
  if ((rc = refit_prepare(ctx, bvh))) return rc;
  return refit_launch(ctx, bvh, box, expand, aos);
}

// Refit only (same leaf order and topology), used when boxes move but the proxy set is unchanged.
int lbvh_refit(ncx_ctx* ctx, LBVH& bvh, const float* const box[6], float expand, const float4* aos) {
  if (bvh.n <= 1) return NCX_OK;
  return refit_launch(ctx, bvh, box, expand, aos);
}
