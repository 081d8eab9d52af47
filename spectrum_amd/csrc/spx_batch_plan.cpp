// Host side of the uneven-length batch: work-list planning for the persistent
// LDS-tile launches (lengths 2^10 .. 2^14, one launch per tile shape in use),
// the split-path launch sequences of the longer lengths (2^15 .. 2^17, one per
// length, in launch groups, channel indirection through the plan's
// permutation and offsets), the autotuner (radix order and tile shape per
// LDS-tile length, launch groups per split length) and its JSON cache, and
// hipGraph capture of the whole sequence.
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

// Tile shapes of the persistent LDS-tile launches: TILE complex points in LDS,
// TILE/16 threads per workgroup, 16384/TILE workgroups per CU.
constexpr int kTiles[3] = {16384, 8192, 4096};

// one persistent launch: every item of one tile shape
struct TileLaunch {
  int tile = 0;
  BatchArgs args{};
  int grid = 0;
  uint32_t n_items = 0;
};

// What the autotuner chooses per length: the radix order and tile shape of an
// LDS-tile length, the number of launch groups a split length's channels are
// divided into (each group one column / row / finish sequence).
struct Choice {
  int order = 0;
  int tile = 16384;
  int groups = 1;
};

struct spx_batch_plan {
  spx_ctx* c = nullptr;
  std::vector<TileLaunch> tl;  // shapes in use, largest tile first
  void* d_items = nullptr;
  void* d_perm = nullptr;
  void* d_offsets = nullptr;
  void* d_counter = nullptr;   // one 64-byte slot per tile launch
  // learned refinement of the LDS-tile channels: their feature rows, the
  // channel lengths (bin widths) and the refiner, run after the tile launches
  void* d_feat = nullptr;
  void* d_lens = nullptr;
  const void* refiner = nullptr;
  uint32_t first_lds = 0, n_lds = 0;
  uint32_t kcap = 0;
  float fs = 1.0f;
  void* d_scratch = nullptr;
  uint64_t scratch_bytes = 0;
  std::vector<SplitGroup> split;
  bool graph = false;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  Choice ch[SPX_BATCH_BUCKETS];
  // the LDS-tile launches on a branch of their own beside the split
  // sequences (a side stream forked from and joined back to the context's
  // stream), instead of after them
  bool concurrent = false;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  uint32_t items_per_bucket[SPX_BATCH_BUCKETS] = {0};
  uint32_t chans_per_bucket[SPX_BATCH_BUCKETS] = {0};
  uint32_t chunk_per_bucket[SPX_BATCH_BUCKETS] = {0};
  uint32_t n_items = 0;
};

namespace {

int log2u(uint32_t n) {
  int l = 0;
  while ((1u << l) < n) ++l;
  return l;
}

uint32_t bucket_len(int bucket) { return 1u << (SPX_BATCH_LOG_MIN + bucket); }
bool is_split(int bucket) { return bucket >= SPX_BATCH_FIRST_LARGE; }

// Length of the transform whose radix order the variant selects: the whole
// channel for LDS tiles, the column transform (L/256) for four-step channels.
uint32_t tuned_len(int bucket) {
  const uint32_t L = bucket_len(bucket);
  return is_split(bucket) ? L / 256 : L;
}

// the split path of the long lengths has one radix order (radix 16 first)
bool has_variants(int bucket) { return !is_split(bucket) && (log2u(tuned_len(bucket)) % 4) != 0; }

bool tile_fits(int bucket, int tile) { return spx_batch_tile_ok(tile) && (uint32_t)tile >= bucket_len(bucket); }

// A choice as the bucket can take it (anything else falls back to the defaults).
Choice sanitize(int bucket, Choice x) {
  Choice r;
  if (is_split(bucket)) {
    r.groups = (x.groups >= 1 && x.groups <= 64) ? x.groups : 1;
  } else {
    r.order = (has_variants(bucket) && x.order == 1) ? 1 : 0;
    r.tile = tile_fits(bucket, x.tile) ? x.tile : 16384;
  }
  return r;
}

// integer value of "key": in s[from, to), or -1
int json_int(const std::string& s, size_t from, size_t to, const char* key) {
  const std::string k = std::string("\"") + key + "\":";
  const size_t p = s.find(k, from);
  if (p == std::string::npos || p >= to) return -1;
  int v = -1;
  return std::sscanf(s.c_str() + p + k.size(), " %d", &v) == 1 ? v : -1;
}

// Cache format (written by spx_batch_autotune):
// {"device": "...", "cu": N, "concurrent": 0|1, "lengths": [{"length": L, "order": V, "tile": T, "groups": g, "us": best,
//   "candidates": [...]}, ...]}; the fields of an entry come before its candidate list, and
// "tile" / "groups" may be absent (defaults).
void read_cache(const char* path, Choice* ch, bool* concurrent) {
  if (!path || !*path) return;
  std::ifstream f(path);
  if (!f) return;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  *concurrent = json_int(s, 0, s.size(), "concurrent") == 1;
  size_t pos = 0;
  while ((pos = s.find("\"length\":", pos)) != std::string::npos) {
    const size_t nxt = s.find("\"length\":", pos + 9);
    size_t end = s.find("\"candidates\"", pos);
    end = std::min(end, nxt);
    if (end == std::string::npos) end = s.size();
    const int L = json_int(s, pos, end, "length");
    if (L > 0) {
      const int b = log2u((uint32_t)L) - SPX_BATCH_LOG_MIN;
      if (b >= 0 && b < SPX_BATCH_BUCKETS && bucket_len(b) == (uint32_t)L) {
        Choice x;
        x.order = json_int(s, pos, end, "order");
        const int t = json_int(s, pos, end, "tile"), g = json_int(s, pos, end, "groups");
        if (t > 0) x.tile = t;
        if (g > 0) x.groups = g;
        ch[b] = sanitize(b, x);
      }
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
  if (d->refine > SPX_REFINE_LEARNED) return fail(c, SPX_EINVAL, "spx_batch: bad refine");
  if (d->refine == SPX_REFINE_LEARNED && !c->refiner)
    return fail(c, SPX_EINVAL, "spx_batch: learned refinement requested but no refiner is set (spx_set_refiner)");
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
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->side) (void)hipStreamDestroy(p->side);
  for (void* q : {p->d_items, p->d_perm, p->d_offsets, p->d_counter, p->d_scratch, p->d_feat, p->d_lens})
    if (q) (void)hipFree(q);
  delete p;
}

// The plan's launches: the split-path sequences of every long length (the
// longest work), then one persistent launch per tile shape, in stream order
// or (concurrent) on the side stream beside the split sequences.
hipError_t launch_all(spx_batch_plan* p, hipStream_t st) {
  const bool conc = p->concurrent && p->side && !p->tl.empty() && !p->split.empty();
  hipStream_t ts = st;
  hipError_t e;
  // the tile launches, then (learned refinement) the refiner over their channels
  auto tiles = [&](hipStream_t s) -> hipError_t {
    for (const TileLaunch& t : p->tl) {
      hipError_t r = spx_launch_batch(&t.args, t.tile, t.grid, s);
      if (r != hipSuccess) return r;
    }
    if (p->d_feat)
      return spx_launch_mlp_refine(p->refiner, p->d_feat, p->tl[0].args.peaks, p->tl[0].args.counts, p->n_lds,
                                   p->kcap, 0.0f, s, (const uint32_t*)p->d_perm + p->first_lds,
                                   (const uint32_t*)p->d_lens, p->fs);
    return hipSuccess;
  };
  if (conc) {
    if ((e = hipEventRecord(p->ev_fork, st)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(p->side, p->ev_fork, 0)) != hipSuccess) return e;
    ts = p->side;
    if ((e = tiles(ts)) != hipSuccess) return e;
  }
  for (const SplitGroup& g : p->split) {
    e = spx_launch_fourstep_ex(&g.fa, p->d_scratch, g.chunk, (const uint32_t*)p->d_perm + g.first,
                               (const uint64_t*)p->d_offsets, st);
    if (e != hipSuccess) return e;
  }
  if (conc) {
    if ((e = hipEventRecord(p->ev_join, p->side)) != hipSuccess) return e;
    return hipStreamWaitEvent(st, p->ev_join, 0);
  }
  return tiles(st);
}

// Builds a plan.  The autotuner's measurement plans pass `force` (a choice
// per bucket and the concurrency, used instead of the cache) and `mask`
// (bit b: keep the channels of bucket b; ~0u keeps all).
int build(spx_ctx* c, const spx_batch_desc* d, uint32_t mask, const Choice* force, bool force_conc, bool graph,
          spx_batch_plan** out) {
  int rc = check_desc(c, d);
  if (rc) return rc;
  spx_batch_plan* p = new spx_batch_plan();
  p->c = c;
  const bool learn = d->refine == SPX_REFINE_LEARNED;
  if (force) {
    for (int b = 0; b < SPX_BATCH_BUCKETS; ++b) p->ch[b] = sanitize(b, force[b]);
    p->concurrent = force_conc;
  } else {
    read_cache(d->tune_cache, p->ch, &p->concurrent);
  }

  // group channels by bucket (ascending channel id inside a bucket)
  std::vector<std::vector<uint32_t>> by(SPX_BATCH_BUCKETS);
  for (uint32_t ch = 0; ch < d->n_chan; ++ch) {
    const int b = log2u(d->lengths[ch]) - SPX_BATCH_LOG_MIN;
    if (!((mask >> b) & 1u)) continue;
    by[b].push_back(ch);
  }
  // long lengths: one split-path group each (their channels first in perm),
  // run in `groups` launch groups; short lengths: one work list of LDS tiles
  // per tile shape, longest-processing-time first
  std::vector<uint32_t> perm;
  std::vector<uint32_t> items[3];  // per tile shape, 4 words per item
  for (int b = SPX_BATCH_BUCKETS - 1; b >= 0; --b) {
    const auto& v = by[b];
    p->chans_per_bucket[b] = (uint32_t)v.size();
    if (v.empty()) continue;
    const uint32_t L = bucket_len(b);
    if (is_split(b)) {
      SplitGroup g;
      std::memset(&g.fa, 0, sizeof(g.fa));
      g.fa.batch = (uint32_t)v.size();
      g.first = (uint32_t)perm.size();
      const uint32_t per_group = (g.fa.batch + (uint32_t)p->ch[b].groups - 1) / (uint32_t)p->ch[b].groups;
      g.chunk = std::min(fourstep_chunk(L, g.fa.batch, learn ? 1 : 0, d->max_peaks), per_group);
      p->split.push_back(g);
      p->items_per_bucket[b] = (g.fa.batch + g.chunk - 1) / g.chunk;
      p->chunk_per_bucket[b] = g.chunk;
    } else {
      const int ti = p->ch[b].tile == 16384 ? 0 : p->ch[b].tile == 8192 ? 1 : 2;
      const uint32_t G = (uint32_t)kTiles[ti] / L;
      const uint32_t word = (uint32_t)b | ((uint32_t)p->ch[b].order << 8);
      for (size_t i = 0; i < v.size(); i += G) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(G, v.size() - i);
        items[ti].insert(items[ti].end(), {word, (uint32_t)(perm.size() + i), cnt, 0u});
        p->items_per_bucket[b]++;
      }
    }
    perm.insert(perm.end(), v.begin(), v.end());
  }
  for (const SplitGroup& g : p->split) p->first_lds += g.fa.batch;
  p->n_lds = (uint32_t)perm.size() - p->first_lds;
  std::vector<uint32_t> all_items;
  std::vector<size_t> item_first;  // first item of each launch in all_items
  for (int ti = 0; ti < 3; ++ti) {
    if (items[ti].empty()) continue;
    TileLaunch t;
    t.tile = kTiles[ti];
    t.n_items = (uint32_t)(items[ti].size() / 4);
    item_first.push_back(all_items.size() / 4);
    all_items.insert(all_items.end(), items[ti].begin(), items[ti].end());
    p->tl.push_back(t);
  }
  p->n_items = (uint32_t)(all_items.size() / 4);
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
  if ((e = hipMalloc(&p->d_items, std::max<size_t>(all_items.size(), 4) * 4)) != hipSuccess)
    return fail_hip(e, "hipMalloc(items)");
  if ((e = hipMalloc(&p->d_perm, perm.size() * 4)) != hipSuccess) return fail_hip(e, "hipMalloc(perm)");
  if ((e = hipMalloc(&p->d_offsets, off_bytes)) != hipSuccess) return fail_hip(e, "hipMalloc(offsets)");
  if ((e = hipMalloc(&p->d_counter, 64 * 3)) != hipSuccess) return fail_hip(e, "hipMalloc(counter)");
  if (!all_items.empty() &&
      (e = hipMemcpy(p->d_items, all_items.data(), all_items.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
    return fail_hip(e, "upload items");
  if ((e = hipMemcpy(p->d_perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
    return fail_hip(e, "upload perm");
  if ((e = hipMemcpy(p->d_offsets, d->offsets, off_bytes, hipMemcpyHostToDevice)) != hipSuccess)
    return fail_hip(e, "upload offsets");

  BatchArgs a{};
  a.in = d->input;
  a.offsets = (const uint64_t*)p->d_offsets;
  a.perm = (const uint32_t*)p->d_perm;
  a.dtype = d->dtype;
  a.kcap = d->max_peaks;
  a.refine = d->refine;
  a.fs = d->sample_rate;
  a.nsigma = d->nsigma;
  a.peaks = d->peaks_out;
  a.counts = d->counts_out;
  a.gmax = d->gmax_out;
  a.stats = d->stats_out;
  if (learn && p->n_lds > 0) {
    p->refiner = c->refiner;
    p->kcap = d->max_peaks;
    p->fs = d->sample_rate;
    if ((e = hipMalloc(&p->d_feat, (size_t)p->n_lds * d->max_peaks * 32)) != hipSuccess)
      return fail_hip(e, "hipMalloc(features)");
    if ((e = hipMalloc(&p->d_lens, (size_t)d->n_chan * 4)) != hipSuccess) return fail_hip(e, "hipMalloc(lengths)");
    if ((e = hipMemcpy(p->d_lens, d->lengths, (size_t)d->n_chan * 4, hipMemcpyHostToDevice)) != hipSuccess)
      return fail_hip(e, "upload lengths");
    a.mlp_feat = p->d_feat;
    a.feat_base = p->first_lds;
  }
  const void* t = nullptr;
  for (int b = 0; b < SPX_BATCH_FIRST_LARGE; ++b) {
    if (by[b].empty()) continue;
    const uint32_t L = bucket_len(b);
    if ((rc = get_table(c, TW, tuned_len(b), (uint32_t)p->ch[b].order, &t))) { release(p); return rc; }
    a.bk[b].tw = (const float2*)t;
    const float* w = nullptr;
    if ((rc = window_ptr(c, d->window, L, &w))) { release(p); return rc; }
    a.bk[b].window = w;
  }
  const int cu = c->cu_count > 0 ? c->cu_count : 256;
  for (size_t i = 0; i < p->tl.size(); ++i) {
    TileLaunch& tl = p->tl[i];
    tl.args = a;
    tl.args.items = (const uint4*)p->d_items + item_first[i];
    tl.args.n_items = tl.n_items;
    tl.args.counter = (uint32_t*)((char*)p->d_counter + 64 * i);
    tl.grid = (int)std::min<uint64_t>((uint64_t)cu * (16384 / tl.tile), tl.n_items);
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
    fa.mlp = learn ? c->refiner : nullptr;
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
    ws = std::max<uint64_t>(ws, spx_fs_workspace_bytes_ex(L, g.chunk, learn ? 1 : 0, d->max_peaks));
  }
  if ((e = spx_batch_kernel_init()) != hipSuccess) return fail_hip(e, "spx_batch_kernel_init");
  if (p->concurrent && !p->tl.empty() && !p->split.empty()) {
    // default priority, like every stream of the library
    if ((e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking)) != hipSuccess)
      return fail_hip(e, "hipStreamCreate(side)");
    if ((e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming)) != hipSuccess)
      return fail_hip(e, "hipEventCreate(fork)");
    if ((e = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming)) != hipSuccess)
      return fail_hip(e, "hipEventCreate(join)");
  }
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
  return build(ctx, desc, ~0u, nullptr, false, !(desc && (desc->flags & SPX_BATCH_NO_GRAPH)), out);
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
  o << "{\"items\": " << p->n_items << ", \"graph\": " << (p->graph ? "true" : "false")
    << ", \"concurrent\": " << (p->side ? "true" : "false")
    << ", \"scratch_bytes\": " << p->scratch_bytes << ", \"launches\": [";
  for (size_t i = 0; i < p->tl.size(); ++i)
    o << (i ? ", " : "") << "{\"tile\": " << p->tl[i].tile << ", \"threads\": " << p->tl[i].tile / 16
      << ", \"grid\": " << p->tl[i].grid << ", \"items\": " << p->tl[i].n_items << "}";
  o << "], \"lengths\": [";
  bool first = true;
  for (int b = 0; b < SPX_BATCH_BUCKETS; ++b) {
    if (!p->chans_per_bucket[b]) continue;
    o << (first ? "" : ", ") << "{\"length\": " << bucket_len(b) << ", \"order\": " << p->ch[b].order
      << ", \"channels\": " << p->chans_per_bucket[b] << ", \"items\": " << p->items_per_bucket[b];
    if (is_split(b))
      o << ", \"path\": \"split\", \"groups\": " << p->items_per_bucket[b] << ", \"chunk\": " << p->chunk_per_bucket[b];
    else
      o << ", \"path\": \"lds-tile\", \"tile\": " << p->ch[b].tile << ", \"threads\": " << p->ch[b].tile / 16
        << ", \"channels_per_tile\": " << (uint32_t)p->ch[b].tile / bucket_len(b);
    o << "}";
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
  SPX_HIP(c, hipSetDevice(c->device));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  auto done = [&](int r) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return r;
  };
  if (he != hipSuccess) return done(hip_fail(c, he, "hipEventCreate"));
  // median of the last 10 of 13 executes of a plan over the buckets in `mask`
  auto time_us = [&](uint32_t mask, const Choice* x, bool conc, double* us) -> int {
    spx_batch_plan* p = nullptr;
    int r = build(c, &dd, mask, x, conc, false, &p);
    if (r) return r;
    std::vector<float> t;
    for (int k = 0; k < 13 && r == SPX_OK; ++k) {
      if (hipEventRecord(e0, c->stream) != hipSuccess) r = SPX_EHIP;
      if (!r) r = run(p);
      if (!r && hipEventRecord(e1, c->stream) != hipSuccess) r = SPX_EHIP;
      if (!r && hipEventSynchronize(e1) != hipSuccess) r = SPX_EHIP;
      float ms = 0;
      if (!r && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) r = SPX_EHIP;
      if (k >= 3) t.push_back(ms);
    }
    release(p);
    if (r) return fail(c, r, "spx_batch_autotune: timing run failed");
    std::sort(t.begin(), t.end());
    *us = 1000.0 * t[t.size() / 2];
    return SPX_OK;
  };
  Choice best[SPX_BATCH_BUCKETS];
  std::vector<std::vector<std::pair<Choice, double>>> cands(SPX_BATCH_BUCKETS);
  uint32_t lds_mask = 0;
  // 1. every length alone: each (radix order, tile shape) of an LDS-tile
  //    length, 1 / 2 / 4 launch groups (of at least 16 channels) of a split one
  for (int b = 0; b < SPX_BATCH_BUCKETS; ++b) {
    if (!count[b]) continue;
    std::vector<Choice> cs;
    if (is_split(b)) {
      for (int g : {1, 2, 4})
        if (g == 1 || count[b] / (uint32_t)g >= 16) {
          Choice x;
          x.groups = g;
          cs.push_back(x);
        }
    } else {
      lds_mask |= 1u << b;
      for (int tile : kTiles)
        for (int v = 0; v < (has_variants(b) ? 2 : 1); ++v)
          if (tile_fits(b, tile)) {
            Choice x;
            x.order = v;
            x.tile = tile;
            cs.push_back(x);
          }
    }
    // the fastest, but the default (the first candidate) unless it is 2 % faster
    double bu = 0;
    for (size_t i = 0; i < cs.size(); ++i) {
      Choice f[SPX_BATCH_BUCKETS];
      f[b] = cs[i];
      double us = 0;
      if ((rc = time_us(1u << b, f, false, &us))) return done(rc);
      cands[b].push_back({cs[i], us});
      if (i == 0 || us < bu) {
        bu = us;
        best[b] = cs[i];
      }
    }
    if (!(bu < 0.98 * cands[b][0].second)) best[b] = cands[b][0].first;
  }
  // 2. the LDS-tile lengths share persistent launches (one per tile shape in
  //    use), so their shapes are chosen together: the assignments "every
  //    length on its largest fitting tile up to 16384 / 8192 / 4096" and the
  //    per-length fastest one, each timed over all LDS-tile lengths at once
  struct TileSet {
    Choice ch[SPX_BATCH_BUCKETS];
    double us = 0;
  };
  std::vector<TileSet> sets;
  if (lds_mask) {
    // radix order on a tile shape: 1 only where it is 2 % faster than 0
    auto order_for = [&](int b, int tile) {
      double t[2] = {-1, -1};
      for (const auto& pc : cands[b])
        if (pc.first.tile == tile) t[pc.first.order] = pc.second;
      return (t[1] > 0 && t[0] > 0 && t[1] < 0.98 * t[0]) ? 1 : 0;
    };
    for (int cap : kTiles) {
      TileSet t;
      for (int b = 0; b < SPX_BATCH_FIRST_LARGE; ++b) {
        if (!count[b]) continue;
        // the largest fitting tile <= cap, else the smallest fitting one
        // (kTiles descend)
        int tile = 0;
        for (int tt : kTiles)
          if (tt <= cap && tile_fits(b, tt)) { tile = tt; break; }
        if (!tile)
          for (int tt : kTiles)
            if (tile_fits(b, tt)) tile = tt;
        t.ch[b].tile = tile;
        t.ch[b].order = order_for(b, tile);
      }
      sets.push_back(t);
    }
    TileSet own;
    for (int b = 0; b < SPX_BATCH_FIRST_LARGE; ++b)
      if (count[b]) own.ch[b] = best[b];
    sets.push_back(own);
    for (TileSet& t : sets)
      if ((rc = time_us(lds_mask, t.ch, false, &t.us))) return done(rc);
    size_t w = 0;
    for (size_t i = 1; i < sets.size(); ++i)
      if (sets[i].us < sets[w].us) w = i;
    for (int b = 0; b < SPX_BATCH_FIRST_LARGE; ++b)
      if (count[b]) best[b] = sets[w].ch[b];
  }
  // 3. with split lengths present too: the whole batch with the LDS-tile
  //    launches after the split sequences or beside them, for every tile set
  //    (beside them, small tiles can share CUs with the split kernels); each
  //    candidate timed in three alternating rounds (median), and the default
  //    (first set, in sequence) kept unless another is 2 % faster
  bool conc = false;
  std::vector<double> whole[2];
  const bool any_split = (lds_mask != 0) && [&] {
    for (int b = SPX_BATCH_FIRST_LARGE; b < SPX_BATCH_BUCKETS; ++b)
      if (count[b]) return true;
    return false;
  }();
  if (any_split) {
    const size_t nc = sets.size() * 2;
    std::vector<std::vector<double>> rounds(nc);
    for (int round = 0; round < 3; ++round)
      for (size_t k = 0; k < nc; ++k) {
        Choice f[SPX_BATCH_BUCKETS];
        for (int b = 0; b < SPX_BATCH_BUCKETS; ++b) f[b] = is_split(b) ? best[b] : sets[k / 2].ch[b];
        double us = 0;
        if ((rc = time_us(~0u, f, (k & 1) == 1, &us))) return done(rc);
        rounds[k].push_back(us);
      }
    std::vector<double> med(nc);
    for (size_t k = 0; k < nc; ++k) {
      std::sort(rounds[k].begin(), rounds[k].end());
      med[k] = rounds[k][1];
      whole[k & 1].push_back(med[k]);
    }
    size_t bk = 0;
    for (size_t k = 1; k < nc; ++k)
      if (med[k] < med[bk]) bk = k;
    if (!(med[bk] < 0.98 * med[0])) bk = 0;
    conc = (bk & 1) == 1;
    for (int b = 0; b < SPX_BATCH_FIRST_LARGE; ++b)
      if (count[b]) best[b] = sets[bk / 2].ch[b];
  }
  std::ostringstream o;
  o << "{\"device\": \"" << c->name << "\", \"cu\": " << c->cu_count << ", \"concurrent\": " << (conc ? 1 : 0)
    << ", \"lengths\": [";
  bool first = true;
  char line[200];
  for (int b = 0; b < SPX_BATCH_BUCKETS; ++b) {
    if (!count[b]) continue;
    const Choice& x = best[b];
    double us = 0;
    for (const auto& pc : cands[b])
      if (pc.first.order == x.order && pc.first.tile == x.tile && pc.first.groups == x.groups) us = pc.second;
    std::snprintf(line, sizeof line, "{\"length\": %u, \"order\": %d, \"tile\": %d, \"groups\": %d, \"us\": %.3f, "
                  "\"channels\": %u, \"candidates\": [", bucket_len(b), x.order, is_split(b) ? 0 : x.tile,
                  is_split(b) ? x.groups : 1, us, count[b]);
    o << (first ? "\n  " : ",\n  ") << line;
    for (size_t i = 0; i < cands[b].size(); ++i) {
      const Choice& y = cands[b][i].first;
      std::snprintf(line, sizeof line, "%s{\"order\": %d, \"tile\": %d, \"groups\": %d, \"us\": %.3f}", i ? ", " : "",
                    y.order, is_split(b) ? 0 : y.tile, is_split(b) ? y.groups : 1, cands[b][i].second);
      o << line;
    }
    o << "]}";
    first = false;
  }
  o << "\n], \"tile_sets\": [";
  for (size_t i = 0; i < sets.size(); ++i) {
    o << (i ? ",\n  " : "\n  ") << "{\"us\": ";
    std::snprintf(line, sizeof line, "%.3f", sets[i].us);
    o << line;
    if (any_split) {
      std::snprintf(line, sizeof line, ", \"whole_us\": %.3f, \"whole_concurrent_us\": %.3f", whole[0][i], whole[1][i]);
      o << line;
    }
    o << ", \"tiles\": {";
    bool f2 = true;
    for (int b = 0; b < SPX_BATCH_FIRST_LARGE; ++b)
      if (count[b]) {
        o << (f2 ? "" : ", ") << "\"" << bucket_len(b) << "\": " << sets[i].ch[b].tile;
        f2 = false;
      }
    o << "}}";
  }
  o << "]}\n";
  std::ofstream f(cache_path);
  if (!f) return done(fail(c, SPX_EINVAL, std::string("spx_batch_autotune: cannot write ") + cache_path));
  f << o.str();
  return done(f.good() ? SPX_OK : fail(c, SPX_EINVAL, "spx_batch_autotune: write failed"));
}

}  // extern "C"
