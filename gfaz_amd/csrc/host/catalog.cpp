// Traversal catalog (see catalog.hpp) and its C ABI. Lookup semantics and error
// messages follow src/workflows/extraction_workflow.cpp:206-424; line text follows
// format_path_line / format_walk_line there (:150-193).
#include "catalog.hpp"

#include "abi_error.hpp"

#include <cstring>
#include <mutex>

namespace gfz {

namespace {

std::string dec(int64_t v) { return v >= 0 ? std::to_string(v) : std::string("*"); }

void put_u64(std::string &s, uint64_t v) { s.append(reinterpret_cast<const char *>(&v), 8); }

}  // namespace

std::string Catalog::tuple_key(const WalkKey &k) {
  std::string s;
  s.reserve(k.sample_id.size() + k.seq_id.size() + 36);
  put_u64(s, k.sample_id.size());
  s += k.sample_id;
  put_u64(s, k.hap_index);
  put_u64(s, k.seq_id.size());
  s += k.seq_id;
  put_u64(s, (uint64_t)k.seq_start);
  put_u64(s, (uint64_t)k.seq_end);
  return s;
}

std::string Catalog::walk_head(const WalkKey &k) {
  return "W\t" + k.sample_id + "\t" + std::to_string(k.hap_index) + "\t" + k.seq_id + "\t" +
         dec(k.seq_start) + "\t" + dec(k.seq_end) + "\t";
}

void Catalog::add_paths(const std::vector<std::string> &names,
                        const std::vector<std::string> &overlaps, uint64_t first_trav_id) {
  if (!overlaps.empty() && overlaps.size() != names.size())
    throw Error(GFZ_E_INVALID, "overlaps must be empty or one per path");
  std::unique_lock<std::shared_mutex> lk(mu_);
  for (size_t i = 0; i < names.size(); ++i) {
    const uint64_t idx = path_names_.size();
    path_names_.push_back(names[i]);
    overlaps_.push_back(overlaps.empty() ? std::string() : overlaps[i]);
    path_ids_.push_back(first_trav_id + i);
    if (!path_index_.emplace(names[i], idx).second && !path_has_repeat_) {
      path_has_repeat_ = true;
      path_repeat_ = names[i];
    }
  }
}

void Catalog::add_walks(const std::vector<WalkKey> &walks, uint64_t first_trav_id) {
  std::unique_lock<std::shared_mutex> lk(mu_);
  for (size_t i = 0; i < walks.size(); ++i) {
    const uint64_t idx = walks_.size();
    walks_.push_back(walks[i]);
    walk_ids_.push_back(first_trav_id + i);
    walk_tuple_.emplace(tuple_key(walks[i]), idx);
    if (!walk_sample_.emplace(walks[i].sample_id, idx).second && !walk_has_repeat_) {
      walk_has_repeat_ = true;
      walk_repeat_ = walks[i].sample_id;
    }
  }
}

uint64_t Catalog::n_paths() const {
  std::shared_lock<std::shared_mutex> lk(mu_);
  return path_names_.size();
}

uint64_t Catalog::n_walks() const {
  std::shared_lock<std::shared_mutex> lk(mu_);
  return walks_.size();
}

std::vector<LineSpec> Catalog::path_lines(const std::vector<std::string> &names) const {
  std::shared_lock<std::shared_mutex> lk(mu_);
  std::vector<LineSpec> out;
  if (names.empty())
    return out;
  if (path_has_repeat_)
    throw Error(GFZ_E_INVALID, "path name is ambiguous: " + path_repeat_);
  out.reserve(names.size());
  for (const std::string &n : names) {
    const auto it = path_index_.find(n);
    if (it == path_index_.end())
      throw Error(GFZ_E_INVALID, "path not found: " + n);
    const std::string &ov = overlaps_[it->second];
    out.push_back({path_ids_[it->second], "P\t" + n + "\t",
                   "\t" + (ov.empty() ? std::string("*") : ov) + "\n"});
  }
  return out;
}

std::vector<LineSpec> Catalog::walk_lines(const std::vector<WalkKey> &keys) const {
  std::shared_lock<std::shared_mutex> lk(mu_);
  std::vector<LineSpec> out;
  out.reserve(keys.size());
  for (const WalkKey &k : keys) {
    const auto it = walk_tuple_.find(tuple_key(k));
    if (it == walk_tuple_.end())
      throw Error(GFZ_E_INVALID, "walk not found for the provided identifier: " + k.sample_id +
                                     "\t" + std::to_string(k.hap_index) + "\t" + k.seq_id +
                                     "\t" + std::to_string(k.seq_start) + "\t" +
                                     std::to_string(k.seq_end));
    out.push_back({walk_ids_[it->second], walk_head(k), "\n"});
  }
  return out;
}

std::vector<LineSpec> Catalog::walk_lines_by_name(const std::vector<std::string> &names) const {
  std::shared_lock<std::shared_mutex> lk(mu_);
  std::vector<LineSpec> out;
  if (names.empty())
    return out;
  if (walk_has_repeat_)
    throw Error(GFZ_E_INVALID, "walk name is ambiguous: " + walk_repeat_);
  out.reserve(names.size());
  for (const std::string &n : names) {
    const auto it = walk_sample_.find(n);
    if (it == walk_sample_.end())
      throw Error(GFZ_E_INVALID, "walk not found: " + n);
    // Sample IDs are unique here, so the stored walk's five fields find it again
    // in extract_walk_lines' scan.
    out.push_back({walk_ids_[it->second], walk_head(walks_[it->second]), "\n"});
  }
  return out;
}

std::vector<std::string> strings_from_abi(const gfz_strings *s, const char *what) {
  std::vector<std::string> out;
  if (!s || !s->n)
    return out;
  if (!s->off || (s->off[s->n] && !s->bytes))
    throw Error(GFZ_E_INVALID, std::string("null ") + what);
  out.reserve(s->n);
  for (uint64_t i = 0; i < s->n; ++i) {
    if (s->off[i + 1] < s->off[i])
      throw Error(GFZ_E_INVALID, std::string(what) + " offsets decrease");
    out.emplace_back(s->bytes + s->off[i], s->off[i + 1] - s->off[i]);
  }
  return out;
}

std::vector<WalkKey> walk_keys_from_abi(const gfz_walk_keys *k) {
  std::vector<WalkKey> out;
  if (!k)
    throw Error(GFZ_E_INVALID, "null walk keys");
  const std::vector<std::string> samples = strings_from_abi(&k->sample_ids, "sample ids");
  const std::vector<std::string> seqs = strings_from_abi(&k->seq_ids, "sequence ids");
  const uint64_t n = samples.size();
  if (seqs.size() != n)
    throw Error(GFZ_E_INVALID, "sample and sequence id counts differ");
  if (n && (!k->hap_indices || !k->seq_starts || !k->seq_ends))
    throw Error(GFZ_E_INVALID, "null walk key column");
  out.resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    out[i].sample_id = samples[i];
    out[i].hap_index = k->hap_indices[i];
    out[i].seq_id = seqs[i];
    out[i].seq_start = k->seq_starts[i];
    out[i].seq_end = k->seq_ends[i];
  }
  return out;
}

}  // namespace gfz

namespace {
constexpr const char *kPrefix = "Extraction workflow error: ";

void copy_ids(const std::vector<gfz::LineSpec> &lines, uint64_t *trav_ids) {
  for (size_t i = 0; i < lines.size(); ++i)
    trav_ids[i] = lines[i].trav_id;
}
}  // namespace

extern "C" {

int gfz_catalog_create(gfz_catalog **out) {
  if (!out)
    return gfz::fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  *out = nullptr;
  return gfz::guarded(kPrefix, [&] { *out = new gfz_catalog(); });
}

void gfz_catalog_free(gfz_catalog *c) { delete c; }

int gfz_catalog_add_paths(gfz_catalog *c, const gfz_strings *names, const gfz_strings *overlaps,
                          uint64_t first_trav_id) {
  if (!c || !names)
    return gfz::fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  return gfz::guarded(kPrefix, [&] {
    c->cat.add_paths(gfz::strings_from_abi(names, "path names"),
                     gfz::strings_from_abi(overlaps, "overlaps"), first_trav_id);
  });
}

int gfz_catalog_add_walks(gfz_catalog *c, const gfz_walk_keys *walks, uint64_t first_trav_id) {
  if (!c || !walks)
    return gfz::fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  return gfz::guarded(kPrefix,
                      [&] { c->cat.add_walks(gfz::walk_keys_from_abi(walks), first_trav_id); });
}

int gfz_catalog_info(const gfz_catalog *c, uint64_t *n_paths, uint64_t *n_walks) {
  if (!c)
    return gfz::fail(GFZ_E_INVALID, std::string(kPrefix) + "null catalog");
  if (n_paths)
    *n_paths = c->cat.n_paths();
  if (n_walks)
    *n_walks = c->cat.n_walks();
  return GFZ_OK;
}

int gfz_catalog_find_paths(const gfz_catalog *c, const gfz_strings *names, uint64_t *trav_ids) {
  if (!c || !names || (names->n && !trav_ids))
    return gfz::fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  return gfz::guarded(kPrefix, [&] {
    copy_ids(c->cat.path_lines(gfz::strings_from_abi(names, "path names")), trav_ids);
  });
}

int gfz_catalog_find_walks(const gfz_catalog *c, const gfz_walk_keys *keys, uint64_t *trav_ids) {
  if (!c || !keys || (keys->sample_ids.n && !trav_ids))
    return gfz::fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  return gfz::guarded(kPrefix, [&] {
    copy_ids(c->cat.walk_lines(gfz::walk_keys_from_abi(keys)), trav_ids);
  });
}

int gfz_catalog_find_walks_by_name(const gfz_catalog *c, const gfz_strings *names,
                                   uint64_t *trav_ids) {
  if (!c || !names || (names->n && !trav_ids))
    return gfz::fail(GFZ_E_INVALID, std::string(kPrefix) + "null argument");
  return gfz::guarded(kPrefix, [&] {
    copy_ids(c->cat.walk_lines_by_name(gfz::strings_from_abi(names, "walk names")), trav_ids);
  });
}

}  // extern "C"
