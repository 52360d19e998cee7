// Seeded synthetic pangenome generator (host only, OpenMP).
//
// Produces GFA path/walk traversals as int32 signed node IDs (sign = orientation,
// IDs dense and 1-based in topological order, so segment names equal IDs and the
// reference parser's numeric fast path applies: src/io/gfa_parser.cpp:605-615,
// 1210-1219), and optionally the GFA text in the direct writer's line format
// (src/io/gfa_writer.cpp:597-640, src/io/gfa_write_utils.cpp:238-290) so that a
// compress/decompress round trip through the reference is byte-identical.
//
// Model (one backbone of `backbone` sites):
//   * site i is a 2-allele bubble when i % 10 == 9; its ref and alt segments get
//     consecutive IDs, so ref(i) = 1 + i + i/10 and alt(i) = ref(i) + 1;
//   * segment lengths are uniform in [1, 64] with i.i.d. ACGT bases;
//   * a haplotype walks the sites; at each site, with p_del it skips U{1..del_max}
//     sites, otherwise it emits the site's segment (bubble allele by a fair coin),
//     reverse-oriented with p_inv; after each emitted step, with p_dup it re-emits
//     the previous U{dup_min..dup_max} steps (a tandem duplication);
//   * P mode: every haplotype walks the whole backbone ("P\thap{h}\t...\t*");
//     W mode: haplotype h targets a log-uniform length in [len_min, len_max] steps,
//     starting at a uniform site ("W\tS{h/2}\t{h%2}\tchr{h}\t{start}\t{end}\t...");
//   * L-lines are the deduplicated, sorted set of adjacencies the haplotypes use,
//     with overlap "*".
// Every random stream is a splitmix64 sequence keyed by (seed, role, index), so
// output is independent of thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

namespace {

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t range(uint64_t lo, uint64_t hi) { return lo + next() % (hi - lo + 1); }
};

uint64_t mix_key(uint64_t seed, uint64_t role, uint64_t idx) {
  SplitMix m(seed ^ (role * 0xD1B54A32D192ED03ull));
  m.s += idx * 0x9E3779B97F4A7C15ull;
  return m.next();
}

}  // namespace

extern "C" {

typedef struct gfz_synth_params {
  uint64_t backbone;
  uint32_t n_haps;
  int32_t walks;
  double p_inv;
  double p_del;
  double p_dup;
  uint32_t del_max;
  uint32_t dup_min;
  uint32_t dup_max;
  uint64_t len_min;
  uint64_t len_max;
  uint64_t seed;
} gfz_synth_params;

uint64_t gfz_synth_num_segments(const gfz_synth_params *p) {
  return p->backbone + p->backbone / 10;
}

static inline int32_t site_ref_id(uint64_t i) { return (int32_t)(1 + i + i / 10); }

static uint32_t segment_length(uint64_t seed, uint64_t id) {
  return (uint32_t)(mix_key(seed, 1, id) % 64) + 1;
}

// Per-traversal start site and target length (W mode) or whole backbone (P mode).
static void trav_window(const gfz_synth_params *p, uint32_t h, uint64_t *start_site,
                        uint64_t *target_len) {
  if (!p->walks) {
    *start_site = 0;
    *target_len = UINT64_MAX;
    return;
  }
  SplitMix r(mix_key(p->seed, 3, h));
  const double lo = std::log((double)p->len_min), hi = std::log((double)p->len_max);
  uint64_t len = (uint64_t)std::llround(std::exp(lo + (hi - lo) * r.unit()));
  len = std::max<uint64_t>(1, std::min(len, p->len_max));
  uint64_t span = p->backbone > len ? p->backbone - len : 0;
  *start_site = span ? r.next() % (span + 1) : 0;
  *target_len = len;
}

static void gen_one(const gfz_synth_params *p, uint32_t h, std::vector<int32_t> &out) {
  out.clear();
  uint64_t site, target;
  trav_window(p, h, &site, &target);
  SplitMix r(mix_key(p->seed, 2, h));
  while (site < p->backbone && out.size() < target) {
    if (r.unit() < p->p_del) {
      site += r.range(1, p->del_max);
      continue;
    }
    int32_t id = site_ref_id(site);
    if (site % 10 == 9 && (r.next() & 1))
      id += 1;
    if (r.unit() < p->p_inv)
      id = -id;
    out.push_back(id);
    if (p->p_dup > 0 && r.unit() < p->p_dup) {
      const uint64_t k = std::min<uint64_t>(r.range(p->dup_min, p->dup_max), out.size());
      const size_t base = out.size() - k;
      for (uint64_t j = 0; j < k && out.size() < target; ++j)
        out.push_back(out[base + j]);
    }
    ++site;
  }
  if (out.size() > target)
    out.resize(target);
}

int gfz_synth_traversals(const gfz_synth_params *p, int32_t **nodes, uint64_t *n_nodes,
                         uint32_t **lengths, uint64_t *n_trav) {
  const uint32_t H = p->n_haps;
  std::vector<uint64_t> lens(H);
  // Pass 1: lengths (generation is cheap; run it twice rather than hold copies).
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1)
    for (int64_t h = 0; h < (int64_t)H; ++h) {
      gen_one(p, (uint32_t)h, buf);
      lens[h] = buf.size();
    }
  }
  std::vector<uint64_t> off(H + 1, 0);
  for (uint32_t h = 0; h < H; ++h) {
    if (lens[h] > UINT32_MAX)
      return -1;
    off[h + 1] = off[h] + lens[h];
  }
  int32_t *nd = (int32_t *)std::malloc(std::max<uint64_t>(1, off[H]) * sizeof(int32_t));
  uint32_t *ln = (uint32_t *)std::malloc(std::max<uint32_t>(1, H) * sizeof(uint32_t));
  if (!nd || !ln) {
    std::free(nd);
    std::free(ln);
    return -2;
  }
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1)
    for (int64_t h = 0; h < (int64_t)H; ++h) {
      gen_one(p, (uint32_t)h, buf);
      std::memcpy(nd + off[h], buf.data(), buf.size() * sizeof(int32_t));
    }
  }
  for (uint32_t h = 0; h < H; ++h)
    ln[h] = (uint32_t)lens[h];
  *nodes = nd;
  *lengths = ln;
  *n_nodes = off[H];
  *n_trav = H;
  return 0;
}

void gfz_synth_free(void *ptr) { std::free(ptr); }

static inline uint64_t adj_key(int32_t a, int32_t b) {
  // from id (31 bits) | from '-' | to id (31 bits) | to '-'
  const uint64_t fa = (uint64_t)(a < 0 ? -a : a), fb = (uint64_t)(b < 0 ? -b : b);
  return (fa << 33) | ((uint64_t)(a < 0) << 32) | (fb << 1) | (uint64_t)(b < 0);
}

// Deduplicated, sorted adjacency keys of all consecutive steps.
static std::vector<uint64_t> collect_links(const int32_t *nodes, const uint32_t *lengths,
                                           uint64_t n_trav) {
  std::vector<uint64_t> off(n_trav + 1, 0);
  for (uint64_t t = 0; t < n_trav; ++t)
    off[t + 1] = off[t] + lengths[t];
  const uint64_t pairs = off[n_trav];
  // Distinct adjacencies are few next to the steps (a node has a handful of
  // successors), so the table starts at 8 slots per node ID and grows x4 when
  // more than half full.
  uint64_t max_id = 0;
#pragma omp parallel for reduction(max : max_id)
  for (int64_t i = 0; i < (int64_t)pairs; ++i) {
    const int64_t v = nodes[i];
    const uint64_t a = (uint64_t)(v < 0 ? -v : v);
    max_id = a > max_id ? a : max_id;
  }
  uint64_t cap = 1024;
  while (cap < std::min(pairs / 4, 8 * max_id) + 1024)
    cap <<= 1;
  for (;;) {
    const uint64_t EMPTY = ~0ull;
    std::vector<std::atomic<uint64_t>> table(cap);
    for (auto &x : table)
      x.store(EMPTY, std::memory_order_relaxed);
    std::atomic<uint64_t> count{0};
    std::atomic<bool> full{false};
    const uint64_t mask = cap - 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t t = 0; t < (int64_t)n_trav; ++t) {
      if (full.load(std::memory_order_relaxed))
        continue;
      for (uint64_t i = off[t] + 1; i < off[t + 1]; ++i) {
        if (full.load(std::memory_order_relaxed))
          break;
        const uint64_t k = adj_key(nodes[i - 1], nodes[i]);
        uint64_t slot = (k * 0x9E3779B97F4A7C15ull) >> 20 & mask;
        for (;;) {
          uint64_t cur = table[slot].load(std::memory_order_relaxed);
          if (cur == k)
            break;
          if (cur == EMPTY) {
            if (table[slot].compare_exchange_strong(cur, k)) {
              if (count.fetch_add(1) + 1 > cap / 2)
                full.store(true);
              break;
            }
            if (cur == k)
              break;
          }
          slot = (slot + 1) & mask;
        }
      }
    }
    if (full.load()) {
      cap <<= 2;
      continue;
    }
    std::vector<uint64_t> keys;
    keys.reserve(count.load());
    for (auto &x : table) {
      const uint64_t v = x.load(std::memory_order_relaxed);
      if (v != EMPTY)
        keys.push_back(v);
    }
    std::sort(keys.begin(), keys.end());
    return keys;
  }
}

static inline unsigned ndigits(uint64_t v) {
  unsigned d = 1;
  while (v >= 10) {
    v /= 10;
    ++d;
  }
  return d;
}

static void walk_fields(const gfz_synth_params *p, uint32_t h, uint32_t len, std::string &s) {
  uint64_t start, target;
  trav_window(p, h, &start, &target);
  char buf[160];
  std::snprintf(buf, sizeof buf, "W\tS%u\t%u\tchr%u\t%llu\t%llu\t", h / 2, h % 2, h,
                (unsigned long long)start, (unsigned long long)(start + len));
  s = buf;
}

static void path_prefix(uint32_t h, std::string &s) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "P\thap%u\t", h);
  s = buf;
}

// Byte size of the GFA text the writer below would produce.
int gfz_synth_gfa_bytes(const gfz_synth_params *p, const int32_t *nodes,
                        const uint32_t *lengths, uint64_t n_trav, uint64_t *bytes) {
  const char *header = "H\tVN:Z:1.1\n";
  uint64_t total = std::strlen(header);
  const uint64_t nseg = gfz_synth_num_segments(p);
  uint64_t sbytes = 0;
#pragma omp parallel for reduction(+ : sbytes)
  for (int64_t id = 1; id <= (int64_t)nseg; ++id)
    sbytes += 2 + ndigits((uint64_t)id) + 1 + segment_length(p->seed, (uint64_t)id) + 1;
  total += sbytes;
  const std::vector<uint64_t> links = collect_links(nodes, lengths, n_trav);
  uint64_t lbytes = 0;
#pragma omp parallel for reduction(+ : lbytes)
  for (int64_t i = 0; i < (int64_t)links.size(); ++i) {
    const uint64_t k = links[i];
    lbytes += 2 + ndigits(k >> 33) + 3 + ndigits((k >> 1) & 0x7FFFFFFFull) + 5;
  }
  total += lbytes;
  std::vector<uint64_t> off(n_trav + 1, 0);
  for (uint64_t t = 0; t < n_trav; ++t)
    off[t + 1] = off[t] + lengths[t];
  uint64_t tbytes = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tbytes)
  for (int64_t t = 0; t < (int64_t)n_trav; ++t) {
    std::string pre;
    uint64_t b = 0;
    for (uint64_t i = off[t]; i < off[t + 1]; ++i) {
      const int32_t v = nodes[i];
      b += ndigits((uint64_t)(v < 0 ? -(int64_t)v : v)) + 1;
    }
    if (p->walks) {
      walk_fields(p, (uint32_t)t, lengths[t], pre);
      b += pre.size() + 1;
    } else {
      path_prefix((uint32_t)t, pre);
      b += pre.size() + (lengths[t] ? lengths[t] - 1 : 0) + 3;
    }
    tbytes += b;
  }
  total += tbytes;
  *bytes = total;
  return 0;
}

int gfz_synth_write_gfa(const gfz_synth_params *p, const int32_t *nodes,
                        const uint32_t *lengths, uint64_t n_trav, const char *path) {
  FILE *f = std::fopen(path, "wb");
  if (!f)
    return -1;
  std::string line;
  std::fputs("H\tVN:Z:1.1\n", f);
  const uint64_t nseg = gfz_synth_num_segments(p);
  static const char ACGT[4] = {'A', 'C', 'G', 'T'};
  for (uint64_t id = 1; id <= nseg; ++id) {
    const uint32_t len = segment_length(p->seed, id);
    SplitMix r(mix_key(p->seed, 4, id));
    line = "S\t" + std::to_string(id) + "\t";
    for (uint32_t j = 0; j < len; ++j)
      line += ACGT[r.next() & 3];
    line += '\n';
    std::fwrite(line.data(), 1, line.size(), f);
  }
  for (uint64_t k : collect_links(nodes, lengths, n_trav)) {
    line = "L\t" + std::to_string(k >> 33) + ((k >> 32 & 1) ? "\t-\t" : "\t+\t") +
           std::to_string((k >> 1) & 0x7FFFFFFFull) + ((k & 1) ? "\t-\t*\n" : "\t+\t*\n");
    std::fwrite(line.data(), 1, line.size(), f);
  }
  uint64_t o = 0;
  for (uint64_t t = 0; t < n_trav; ++t) {
    if (p->walks) {
      walk_fields(p, (uint32_t)t, lengths[t], line);
      for (uint32_t i = 0; i < lengths[t]; ++i) {
        const int32_t v = nodes[o + i];
        line += v < 0 ? '<' : '>';
        line += std::to_string(v < 0 ? -(int64_t)v : v);
      }
    } else {
      path_prefix((uint32_t)t, line);
      for (uint32_t i = 0; i < lengths[t]; ++i) {
        const int32_t v = nodes[o + i];
        if (i)
          line += ',';
        line += std::to_string(v < 0 ? -(int64_t)v : v);
        line += v < 0 ? '-' : '+';
      }
      line += "\t*";
    }
    line += '\n';
    std::fwrite(line.data(), 1, line.size(), f);
    o += lengths[t];
  }
  return std::fclose(f) == 0 ? 0 : -2;
}

}  // extern "C"
