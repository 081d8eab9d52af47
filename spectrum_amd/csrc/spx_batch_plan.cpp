// Host side of the uneven-length batch: work-list planning for the persistent
// LDS-tile launch (lengths 2^10 .. 2^14), the split-path launch sequences of
// the longer lengths (2^15 .. 2^17, one per length, channel indirection
// through the plan's permutation and offsets), the radix-order autotuner and
// its JSON cache, and hipGraph capture of the whole sequence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "spx_host.h"

using namespace spx_host;

// one length of the batch on the split path
struct SplitGroup {
  FusedArgs fa;      // batch = count, in = the packed input, tables of the length
  uint32_t first;    // its channels are perm[first .. first + count)
  uint32_t chunk;    // channels per launch sequence
};

struct spx_batch_plan {
  spx_ctx* c = nullptr;
  BatchArgs args{};
  int grid = 0;
  void* d_items = nullptr;
  void* d_perm = nullptr;
  void* d_offsets = nullptr;
  void* d_counter = nullptr;
  void* d_scratch = nullptr;
  uint64_t scratch_bytes = 0;
  std::vector<SplitGroup> split;
  bool graph = false;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  int variant[SPX_BATCH_BUCKETS] = {0};
  uint32_t items_per_bucket[SPX_BATCH_BUCKETS] = {0};
  uint32_t chans_per_bucket[SPX_BATCH_BUCKETS] = {0};
  uint32_t n_items = 0;
};

namespace {

int log2u(uint32_t n) {
  int l = 0;
  while ((1u << l) < n) ++l;
  return l;
}

// Length of the transform whose radix order the variant selects: the whole
// channel for LDS tiles, the column transform (L/256) for four-step channels.
uint32_t tuned_len(int bucket) {
  const uint32_t L = 1u << (SPX_BATCH_LOG_MIN + bucket);
  return bucket >= SPX_BATCH_FIRST_LARGE ? L / 256 : L;
}

// the split path of the long lengths has one radix order (radix 16 first)
bool has_variants(int bucket) { return bucket < SPX_BATCH_FIRST_LARGE && (log2u(tuned_len(bucket)) % 4) != 0; }

// Cache format (written by spx_batch_autotune):
// {"device": "...", "cu": N, "lengths": [{"length": L, "order": V, "us": [t0, t1], "channels": n}, ...]}
void read_cache(const char* path, int* variant) {
  if (!path || !*path) return;
  std::ifstream f(path);
  if (!f) return;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  size_t pos = 0;
  while ((pos = s.find("\"length\":", pos)) != std::string::npos) {
    unsigned L = 0;
    int V = 0;
    if (std::sscanf(s.c_str() + pos, "\"length\": %u, \"order\": %d", &L, &V) == 2) {
      const int b = log2u(L) - SPX_BATCH_LOG_MIN;
      if (b >= 0 && b < SPX_BATCH_BUCKETS && (1u << (SPX_BATCH_LOG_MIN + b)) == L && (V == 0 || V == 1))
        variant[b] = has_variants(b) ? V : 0;
    }
    pos += 9;
  }
}

int check_desc(spx_ctx* c, const spx_batch_desc* d) {
  if (!d || !d->input || !d->offsets || !d->lengths || !d->peaks_out || !d->counts_out)
    return fail(c, SPX_EINVAL, "spx_batch: null pointer in descriptor");
  if (d->n_chan == 0) return fail(c, SPX_EINVAL, "spx_batch: n_chan must be > 0");
  if (d->dtype > SPX_DTYPE_C32BF16) return fail(c, SPX_EINVAL, "spx_batch: bad dtype");
  if (d->window > SPX_WINDOW_BLACKMAN) return fail(c, SPX_EINVAL, "spx_batch: bad window");
  if (d->refine > SPX_REFINE_PARABOLIC) return fail(c, SPX_ENOTSUP, "spx_batch: refine must be NONE or PARABOLIC");
  if (d->max_peaks == 0) return fail(c, SPX_EINVAL, "spx_batch: max_peaks must be >= 1");
  if (!(d->nsigma >= 0.0f)) return fail(c, SPX_EINVAL, "spx_batch: nsigma must be >= 0");
  for (uint32_t i = 0; i < d->n_chan; ++i) {
    const uint32_t L = d->lengths[i];
    if (L < (1u << SPX_BATCH_LOG_MIN) || L > (1u << (SPX_BATCH_LOG_MIN + SPX_BATCH_BUCKETS - 1)) || (L & (L - 1)))
      return fail(c, SPX_EINVAL, "spx_batch: channel " + std::to_string(i) + " length " + std::to_string(L) +
                                     " is not a power of two in [1024, 131072]");
  }
  return SPX_OK;
}

void release(spx_batch_plan* p) {
  if (!p) return;
  if (p->c) (void)hipSetDevice(p->c->device);
  if (p->ge) (void)hipGraphExecDestroy(p->ge);
  if (p->g) (void)hipGraphDestroy(p->g);
  for (void* q : {p->d_items, p->d_perm, p->d_offsets, p->d_counter, p->d_scratch})
    if (q) (void)hipFree(q);
  delete p;
}

// The plan's launches in stream order: the split-path sequence of every long
// length (the longest work), then the persistent LDS-tile launch.
hipError_t launch_all(spx_batch_plan* p, hipStream_t st) {
  for (const SplitGroup& g : p->split) {
    hipError_t e = spx_launch_fourstep_ex(&g.fa, p->d_scratch, g.chunk, (const uint32_t*)p->d_perm + g.first,
                                          (const uint64_t*)p->d_offsets, st);
    if (e != hipSuccess) return e;
  }
  if (p->n_items) return spx_launch_batch(&p->args, p->grid, st);
  return hipSuccess;
}

// Builds a plan; bucket_only >= 0 restricts it to one length and forces
// `force_variant` there (the autotuner's measurement plans).
int build(spx_ctx* c, const spx_batch_desc* d, int bucket_only, int force_variant, bool graph, spx_batch_plan** out) {
  int rc = check_desc(c, d);
  if (rc) return rc;
  spx_batch_plan* p = new spx_batch_plan();
  p->c = c;
  read_cache(d->tune_cache, p->variant);
  if (bucket_only >= 0) p->variant[bucket_only] = has_variants(bucket_only) ? force_variant : 0;

  // group channels by bucket (ascending channel id inside a bucket)
  std::vector<std::vector<uint32_t>> by(SPX_BATCH_BUCKETS);
  for (uint32_t ch = 0; ch < d->n_chan; ++ch) {
    const int b = log2u(d->lengths[ch]) - SPX_BATCH_LOG_MIN;
    if (bucket_only >= 0 && b != bucket_only) continue;
    by[b].push_back(ch);
  }
  // long lengths: one split-path group each (their channels first in perm);
  // short lengths: work list of LDS tiles, longest-processing-time first
  std::vector<uint32_t> perm;
  std::vector<uint32_t> items;  // 4 words per item
  for (int b = SPX_BATCH_BUCKETS - 1; b >= 0; --b) {
    const auto& v = by[b];
    p->chans_per_bucket[b] = (uint32_t)v.size();
    if (v.empty()) continue;
    const uint32_t L = 1u << (SPX_BATCH_LOG_MIN + b);
    if (b >= SPX_BATCH_FIRST_LARGE) {
      SplitGroup g;
      std::memset(&g.fa, 0, sizeof(g.fa));
      g.fa.batch = (uint32_t)v.size();
      g.first = (uint32_t)perm.size();
      g.chunk = fourstep_chunk(L, g.fa.batch);
      p->split.push_back(g);
      p->items_per_bucket[b] = (g.fa.batch + g.chunk - 1) / g.chunk;
    } else {
      const uint32_t G = 16384u / L;
      const uint32_t word = (uint32_t)b | ((uint32_t)p->variant[b] << 8);
      for (size_t i = 0; i < v.size(); i += G) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(G, v.size() - i);
        items.insert(items.end(), {word, (uint32_t)(perm.size() + i), cnt, 0u});
        p->items_per_bucket[b]++;
      }
    }
    perm.insert(perm.end(), v.begin(), v.end());
  }
  p->n_items = (uint32_t)(items.size() / 4);
  if (p->n_items == 0 && p->split.empty()) {
    release(p);
    return fail(c, SPX_EINVAL, "spx_batch: no channels in the requested length");
  }

  auto fail_hip = [&](hipError_t e, const char* what) {
    release(p);
    return hip_fail(c, e, what);
  };
  hipError_t e;
  if ((e = hipSetDevice(c->device)) != hipSuccess) return fail_hip(e, "hipSetDevice");
  const size_t off_bytes = (size_t)d->n_chan * sizeof(uint64_t);
  if ((e = hipMalloc(&p->d_items, std::max<size_t>(items.size(), 4) * 4)) != hipSuccess)
    return fail_hip(e, "hipMalloc(items)");
  if ((e = hipMalloc(&p->d_perm, perm.size() * 4)) != hipSuccess) return fail_hip(e, "hipMalloc(perm)");
  if ((e = hipMalloc(&p->d_offsets, off_bytes)) != hipSuccess) return fail_hip(e, "hipMalloc(offsets)");
  if ((e = hipMalloc(&p->d_counter, 256)) != hipSuccess) return fail_hip(e, "hipMalloc(counter)");
  if (!items.empty() && (e = hipMemcpy(p->d_items, items.data(), items.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
    return fail_hip(e, "upload items");
  if ((e = hipMemcpy(p->d_perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
    return fail_hip(e, "upload perm");
  if ((e = hipMemcpy(p->d_offsets, d->offsets, off_bytes, hipMemcpyHostToDevice)) != hipSuccess)
    return fail_hip(e, "upload offsets");

  BatchArgs& a = p->args;
  a.in = d->input;
  a.offsets = (const uint64_t*)p->d_offsets;
  a.perm = (const uint32_t*)p->d_perm;
  a.items = (const uint4*)p->d_items;
  a.counter = (uint32_t*)p->d_counter;
  a.n_items = p->n_items;
  a.dtype = d->dtype;
  a.kcap = d->max_peaks;
  a.refine = d->refine;
  a.fs = d->sample_rate;
  a.nsigma = d->nsigma;
  a.peaks = d->peaks_out;
  a.counts = d->counts_out;
  a.gmax = d->gmax_out;
  a.stats = d->stats_out;
  const void* t = nullptr;
  for (int b = 0; b < SPX_BATCH_FIRST_LARGE; ++b) {
    if (by[b].empty()) continue;
    const uint32_t L = 1u << (SPX_BATCH_LOG_MIN + b);
    if ((rc = get_table(c, TW, tuned_len(b), (uint32_t)p->variant[b], &t))) { release(p); return rc; }
    a.bk[b].tw = (const float2*)t;
    const float* w = nullptr;
    if ((rc = window_ptr(c, d->window, L, &w))) { release(p); return rc; }
    a.bk[b].window = w;
  }
  uint64_t ws = 0;
  for (SplitGroup& g : p->split) {
    const uint32_t L = d->lengths[perm[g.first]];
    FusedArgs& fa = g.fa;
    fa.in = d->input;
    fa.n_point = L;
    fa.nfft = L;
    fa.dtype = d->dtype;
    if ((rc = window_ptr(c, d->window, L, &fa.window))) { release(p); return rc; }
    fa.fs = d->sample_rate;
    fa.mag_scale = 1.0f;
    fa.peak_mode = 1;
    fa.nsigma = d->nsigma;
    fa.kcap = d->max_peaks;
    fa.s0 = 0;
    fa.s1 = L;
    fa.refine = d->refine;
    fa.peaks = d->peaks_out;
    fa.counts = d->counts_out;
    fa.gmax = d->gmax_out;
    fa.stats = d->stats_out;
    if ((rc = get_table(c, TW, 256, 0, &t))) { release(p); return rc; }
    fa.tw_a = (const float2*)t;
    if ((rc = get_table(c, TW, L / 256, 0, &t))) { release(p); return rc; }
    fa.tw_b = (const float2*)t;
    if ((rc = get_table(c, TW4, L, 0, &t))) { release(p); return rc; }
    fa.tw_4 = (const float2*)t;
    ws = std::max<uint64_t>(ws, spx_fs_workspace_bytes(L, g.chunk));
  }
  if ((e = spx_batch_kernel_init()) != hipSuccess) return fail_hip(e, "spx_batch_kernel_init");
  p->grid = c->cu_count > 0 ? c->cu_count : 256;
  if ((uint32_t)p->grid > p->n_items) p->grid = (int)p->n_items;
  if (ws) {
    p->scratch_bytes = ws;
    if ((e = hipMalloc(&p->d_scratch, p->scratch_bytes)) != hipSuccess) {
      release(p);
      return oom(c, "spx_batch: workspace allocation failed");
    }
  }
  if (graph) {
    if ((e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return fail_hip(e, "hipStreamBeginCapture");
    hipError_t le = launch_all(p, c->stream);
    e = hipStreamEndCapture(c->stream, &p->g);
    if (le != hipSuccess) return fail_hip(le, "batch launches (capture)");
    if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture");
    if ((e = hipGraphInstantiate(&p->ge, p->g, nullptr, nullptr, 0)) != hipSuccess)
      return fail_hip(e, "hipGraphInstantiate");
    p->graph = true;
  }
  *out = p;
  return SPX_OK;
}

int run(spx_batch_plan* p) {
  spx_ctx* c = p->c;
  SPX_HIP(c, hipSetDevice(c->device));
  if (p->graph) {
    SPX_HIP(c, hipGraphLaunch(p->ge, c->stream));
  } else {
    SPX_HIP(c, launch_all(p, c->stream));
  }
  return SPX_OK;
}

}  // namespace

extern "C" {

int spx_batch_plan_create(spx_ctx* ctx, const spx_batch_desc* desc, spx_batch_plan** out) {
  if (!ctx || !out) return SPX_EINVAL;
  *out = nullptr;
  return build(ctx, desc, -1, 0, !(desc && (desc->flags & SPX_BATCH_NO_GRAPH)), out);
}

int spx_batch_execute(spx_batch_plan* plan) {
  if (!plan) return SPX_EINVAL;
  return run(plan);
}

int spx_batch_plan_destroy(spx_batch_plan* plan) {
  if (!plan) return SPX_EINVAL;
  release(plan);
  return SPX_OK;
}

int spx_batch_plan_info(spx_batch_plan* p, char* buf, size_t len) {
  if (!p || !buf || len == 0) return SPX_EINVAL;
  std::ostringstream o;
  o << "{\"items\": " << p->n_items << ", \"grid\": " << p->grid << ", \"graph\": " << (p->graph ? "true" : "false")
    << ", \"scratch_bytes\": " << p->scratch_bytes << ", \"lengths\": [";
  bool first = true;
  for (int b = 0; b < SPX_BATCH_BUCKETS; ++b) {
    if (!p->chans_per_bucket[b]) continue;
    o << (first ? "" : ", ") << "{\"length\": " << (1u << (SPX_BATCH_LOG_MIN + b)) << ", \"order\": " << p->variant[b]
      << ", \"channels\": " << p->chans_per_bucket[b] << ", \"items\": " << p->items_per_bucket[b]
      << ", \"path\": \"" << (b >= SPX_BATCH_FIRST_LARGE ? "split" : "lds-tile") << "\"}";
    first = false;
  }
  o << "]}";
  const std::string s = o.str();
  if (s.size() + 1 > len) return SPX_ECAP;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return SPX_OK;
}

int spx_batch_autotune(spx_ctx* c, const spx_batch_desc* d, const char* cache_path) {
  if (!c || !cache_path) return SPX_EINVAL;
  int rc = check_desc(c, d);
  if (rc) return rc;
  std::vector<uint32_t> count(SPX_BATCH_BUCKETS, 0);
  for (uint32_t i = 0; i < d->n_chan; ++i) count[log2u(d->lengths[i]) - SPX_BATCH_LOG_MIN]++;
  spx_batch_desc dd = *d;
  dd.tune_cache = nullptr;
  hipEvent_t e0, e1;
  SPX_HIP(c, hipSetDevice(c->device));
  SPX_HIP(c, hipEventCreate(&e0));
  SPX_HIP(c, hipEventCreate(&e1));
  std::ostringstream o;
  o << "{\"device\": \"" << c->name << "\", \"cu\": " << c->cu_count << ", \"lengths\": [";
  bool first = true;
  for (int b = 0; b < SPX_BATCH_BUCKETS; ++b) {
    if (!count[b]) continue;
    const int nv = has_variants(b) ? 2 : 1;
    double us[2] = {0, 0};
    for (int v = 0; v < nv; ++v) {
      spx_batch_plan* p = nullptr;
      if ((rc = build(c, &dd, b, v, false, &p))) return rc;
      std::vector<float> t;
      for (int r = 0; r < 7 && rc == SPX_OK; ++r) {
        if (hipEventRecord(e0, c->stream) != hipSuccess) rc = SPX_EHIP;
        if (!rc) rc = run(p);
        if (!rc && hipEventRecord(e1, c->stream) != hipSuccess) rc = SPX_EHIP;
        if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = SPX_EHIP;
        float ms = 0;
        if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = SPX_EHIP;
        if (r >= 2) t.push_back(ms);
      }
      release(p);
      if (rc) return fail(c, rc, "spx_batch_autotune: timing run failed");
      std::sort(t.begin(), t.end());
      us[v] = 1000.0 * t[t.size() / 2];
    }
    const int best = (nv == 2 && us[1] < us[0]) ? 1 : 0;
    char line[256];
    if (nv == 2)
      std::snprintf(line, sizeof line, "{\"length\": %u, \"order\": %d, \"us\": [%.3f, %.3f], \"channels\": %u}",
                    1u << (SPX_BATCH_LOG_MIN + b), best, us[0], us[1], count[b]);
    else
      std::snprintf(line, sizeof line, "{\"length\": %u, \"order\": 0, \"us\": [%.3f], \"channels\": %u}",
                    1u << (SPX_BATCH_LOG_MIN + b), us[0], count[b]);
    const std::string ls(line);
    o << (first ? "\n  " : ",\n  ") << ls;
    first = false;
  }
  o << "\n]}\n";
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::ofstream f(cache_path);
  if (!f) return fail(c, SPX_EINVAL, std::string("spx_batch_autotune: cannot write ") + cache_path);
  f << o.str();
  return f.good() ? SPX_OK : fail(c, SPX_EINVAL, "spx_batch_autotune: write failed");
}

}  // extern "C"
