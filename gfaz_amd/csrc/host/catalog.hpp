// Traversal catalog: path names and walk identifiers of resident traversals,
// indexed once so that a lookup costs a hash probe instead of the reference's
// per-call decode of the name columns and linear tuple scan
// (src/workflows/extraction_workflow.cpp:206-424).
#pragma once

#include "gfaz_amd.h"

#include <cstdint>
#include <shared_mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace gfz {

struct WalkKey {
  std::string sample_id;
  uint32_t hap_index = 0;
  std::string seq_id;
  int64_t seq_start = -1, seq_end = -1;
};

// One line to produce: the traversal to decode and the text around its steps.
struct LineSpec {
  uint64_t trav_id;
  std::string head, tail;
};

class Catalog {
 public:
  // Paths and walks get the traversal IDs first_trav_id, first_trav_id + 1, ...
  // (the resident store's indices); overlaps may be empty (every line "*").
  void add_paths(const std::vector<std::string> &names, const std::vector<std::string> &overlaps,
                 uint64_t first_trav_id);
  void add_walks(const std::vector<WalkKey> &walks, uint64_t first_trav_id);
  uint64_t n_paths() const;
  uint64_t n_walks() const;

  // extract_path_lines_by_name: every stored name must be unique ("path name is
  // ambiguous: <first repeat>"), every requested one present ("path not found").
  std::vector<LineSpec> path_lines(const std::vector<std::string> &names) const;
  // extract_walk_lines: the first stored walk equal in all five fields.
  std::vector<LineSpec> walk_lines(const std::vector<WalkKey> &keys) const;
  // extract_walk_lines_by_name: sample IDs as unique names, the rest from storage.
  std::vector<LineSpec> walk_lines_by_name(const std::vector<std::string> &names) const;

 private:
  static std::string tuple_key(const WalkKey &k);
  static std::string walk_head(const WalkKey &k);

  mutable std::shared_mutex mu_;
  std::vector<std::string> path_names_, overlaps_;
  std::vector<uint64_t> path_ids_;
  std::unordered_map<std::string, uint64_t> path_index_;  // name -> path index (first)
  std::string path_repeat_;                               // first repeated name
  bool path_has_repeat_ = false;

  std::vector<WalkKey> walks_;
  std::vector<uint64_t> walk_ids_;
  std::unordered_map<std::string, uint64_t> walk_tuple_;   // five fields -> first walk
  std::unordered_map<std::string, uint64_t> walk_sample_;  // sample id -> walk (first)
  std::string walk_repeat_;
  bool walk_has_repeat_ = false;
};

// Strings of the C ABI (concatenated bytes, n + 1 offsets) to std::string.
std::vector<std::string> strings_from_abi(const gfz_strings *s, const char *what);
std::vector<WalkKey> walk_keys_from_abi(const gfz_walk_keys *k);

}  // namespace gfz

struct gfz_catalog {
  gfz::Catalog cat;
};
