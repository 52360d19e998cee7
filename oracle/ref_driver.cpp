// Test-infrastructure driver over the reference CPU library (compiled from the
// reference's own sources by oracle/Makefile into oracle/_ref/). It is the
// checker for golden fixtures and the "reference" CPU baseline; nothing in the
// product links or runs it.
//
// Commands (all single-process, OpenMP threads set by -j):
//   compress   <in.gfa> <out.gfaz> <rounds> <threshold> <delta_round> <threads>
//       -> compress_gfa (src/workflows/compression_workflow.cpp:706-825)
//          + serialize_compressed_data (src/codec/serialization.cpp:137-236)
//   decompress <in.gfaz> <out.gfa> <threads>
//       -> deserialize + write_gfa_from_compressed_data (src/io/gfa_writer.cpp:376)
//   extract-path <in.gfaz> <name>...
//       -> extract_path_lines_by_name (src/workflows/extraction_workflow.cpp:206)
//   extract-walk <in.gfaz> <sample> <hap> <seq> <start|*> <end|*>
//       -> extract_walk_line (src/workflows/extraction_workflow.cpp:266)
//   add-haplotypes <in.gfaz> <append.gfa> <out.gfaz>
//       -> add_haplotypes (src/workflows/add_haplotypes_workflow.cpp:579)
//   codec-times <in.gfa> <out.gfaz> <rounds> <threshold> <delta_round> <threads>
//       -> one timed compress_gfa (then serialised, untimed), one timed
//          write_gfa_from_compressed_data to /dev/null (the CLI's decompress) and
//          one timed decompress_gfa (src/workflows/decompression_workflow.cpp:254,
//          the in-memory decoder), printed as JSON. With GFA_COMPRESSION_DEBUG=1
//          in the environment the reference prints its own stage timers to stderr
//          (compression_workflow.cpp:695-698,791-819; decompression_workflow.cpp:
//          260-546), which give the codec-only times.
//   growth <in.gfaz> <mode> <threads>
//       -> compute_growth (src/workflows/growth_workflow.cpp:261), printed as JSON
//          (hist, growth with 17 significant digits); mode 0 PerPathWalk,
//          1 SampleHapSeq, 2 SampleHap, 3 Sample
//   pav <in.gfaz> <in.bed> <mode> <threads>
//       -> compute_pav (src/workflows/pav_workflow.cpp:475), printed as JSON
//          (group names, denominators, numerators)
//   time-compress <in.gfa> <rounds> <threshold> <delta_round> <threads> <reps>
//   time-decompress <in.gfaz> <threads> <reps>
//       -> wall-clock timing of the two calls above, printed as JSON; the
//          decompress timing writes the GFA to /dev/null.
#include "codec/serialization.hpp"
#include "io/gfa_writer.hpp"
#include "workflows/add_haplotypes_workflow.hpp"
#include "workflows/compression_workflow.hpp"
#include "workflows/decompression_workflow.hpp"
#include "workflows/growth_workflow.hpp"
#include "workflows/pav_workflow.hpp"
#include "workflows/extraction_workflow.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

static int64_t parse_coord(const char *s) {
  if (std::strcmp(s, "*") == 0)
    return -1;
  return std::strtoll(s, nullptr, 10);
}

static double now_s() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "usage: gfaz_ref <command> ..." << std::endl;
    return 2;
  }
  const std::string cmd = argv[1];
  try {
    if (cmd == "compress" && argc == 8) {
      gfaz::CompressedData d =
          compress_gfa(argv[2], std::atoi(argv[4]), std::strtoul(argv[5], nullptr, 10),
                       std::atoi(argv[6]), std::atoi(argv[7]), false);
      gfaz::serialize_compressed_data(d, argv[3]);
      return 0;
    }
    if (cmd == "decompress" && argc == 5) {
      gfaz::CompressedData d = gfaz::deserialize_compressed_data(argv[2]);
      write_gfa_from_compressed_data(d, argv[3], std::atoi(argv[4]));
      return 0;
    }
    if (cmd == "extract-path" && argc >= 4) {
      gfaz::CompressedData d = gfaz::deserialize_compressed_data(argv[2]);
      std::vector<std::string> names(argv + 3, argv + argc);
      for (const auto &line : extract_path_lines_by_name(d, names, 1))
        std::cout << line;
      return 0;
    }
    if (cmd == "extract-walk" && argc == 8) {
      gfaz::CompressedData d = gfaz::deserialize_compressed_data(argv[2]);
      std::cout << extract_walk_line(d, argv[3],
                                     static_cast<uint32_t>(std::strtoul(argv[4], nullptr, 10)),
                                     argv[5], parse_coord(argv[6]), parse_coord(argv[7]), 1);
      return 0;
    }
    if (cmd == "add-haplotypes" && argc == 5) {
      gfaz::CompressedData d = gfaz::deserialize_compressed_data(argv[2]);
      add_haplotypes(d, argv[3], 1);
      gfaz::serialize_compressed_data(d, argv[4]);
      return 0;
    }
    if (cmd == "codec-times" && argc == 8) {
      const int threads = std::atoi(argv[7]);
      double t0 = now_s();
      gfaz::CompressedData d =
          compress_gfa(argv[2], std::atoi(argv[4]), std::strtoul(argv[5], nullptr, 10),
                       std::atoi(argv[6]), threads, false);
      const double tc = now_s() - t0;
      gfaz::serialize_compressed_data(d, argv[3]);
      t0 = now_s();
      write_gfa_from_compressed_data(d, "/dev/null", threads);
      const double tw = now_s() - t0;
      double tl;
      {
        gfaz::GfaGraph g;
        t0 = now_s();
        decompress_gfa(d, g, threads);
        tl = now_s() - t0;
      }
      std::printf("{\"compress_s\": %.6f, \"decompress_s\": %.6f, "
                  "\"decompress_in_memory_s\": %.6f, \"threads\": %d}\n",
                  tc, tw, tl, threads);
      return 0;
    }
    if ((cmd == "growth" && argc == 5) || (cmd == "pav" && argc == 6)) {
      static const gfaz::GroupingMode modes[] = {
          gfaz::GroupingMode::PerPathWalk, gfaz::GroupingMode::SampleHapSeq,
          gfaz::GroupingMode::SampleHap, gfaz::GroupingMode::Sample};
      gfaz::CompressedData d = gfaz::deserialize_compressed_data(argv[2]);
      const int m = std::atoi(argv[cmd == "growth" ? 3 : 4]);
      const int threads = std::atoi(argv[cmd == "growth" ? 4 : 5]);
      if (m < 0 || m > 3)
        throw std::runtime_error("mode must be 0..3");
      if (cmd == "growth") {
        gfaz::GrowthResult r = gfaz::compute_growth(d, threads, modes[m]);
        std::printf("{\"num_haplotypes\": %u, \"num_nodes\": %u, \"hist\": [", r.num_haplotypes,
                    r.num_nodes);
        for (size_t i = 0; i < r.hist.size(); ++i)
          std::printf("%s%llu", i ? ", " : "", (unsigned long long)r.hist[i]);
        std::printf("], \"growth\": [");
        for (size_t i = 0; i < r.growth.size(); ++i)
          std::printf("%s%.17g", i ? ", " : "", r.growth[i]);
        std::printf("]}\n");
      } else {
        gfaz::PavOptions o;
        o.bed_path = argv[3];
        o.num_threads = threads;
        o.grouping = modes[m];
        gfaz::PavResult r = gfaz::compute_pav(d, o);
        std::printf("{\"groups\": [");
        for (size_t i = 0; i < r.group_names.size(); ++i)
          std::printf("%s\"%s\"", i ? ", " : "", r.group_names[i].c_str());
        std::printf("], \"ranges\": [");
        for (size_t i = 0; i < r.ranges.size(); ++i)
          std::printf("%s[\"%s\", %llu, %llu, \"%s\"]", i ? ", " : "", r.ranges[i].chrom.c_str(),
                      (unsigned long long)r.ranges[i].start, (unsigned long long)r.ranges[i].end,
                      r.ranges[i].name.c_str());
        std::printf("], \"denominators\": [");
        for (size_t i = 0; i < r.denominators.size(); ++i)
          std::printf("%s%llu", i ? ", " : "", (unsigned long long)r.denominators[i]);
        std::printf("], \"numerators\": [");
        for (size_t i = 0; i < r.numerators.size(); ++i)
          std::printf("%s%llu", i ? ", " : "", (unsigned long long)r.numerators[i]);
        std::printf("]}\n");
      }
      return 0;
    }
    if (cmd == "time-compress" && argc == 8) {
      const int reps = std::max(1, std::atoi(argv[7]));
      double best = 1e30;
      for (int r = 0; r < reps; ++r) {
        const double t0 = now_s();
        gfaz::CompressedData d =
            compress_gfa(argv[2], std::atoi(argv[3]), std::strtoul(argv[4], nullptr, 10),
                         std::atoi(argv[5]), std::atoi(argv[6]), false);
        const double t1 = now_s();
        if (t1 - t0 < best)
          best = t1 - t0;
      }
      std::printf("{\"op\": \"compress\", \"seconds\": %.6f, \"threads\": %d}\n", best,
                  std::atoi(argv[6]));
      return 0;
    }
    if (cmd == "time-decompress" && argc == 5) {
      const int reps = std::max(1, std::atoi(argv[4]));
      gfaz::CompressedData d = gfaz::deserialize_compressed_data(argv[2]);
      double best = 1e30;
      for (int r = 0; r < reps; ++r) {
        const double t0 = now_s();
        write_gfa_from_compressed_data(d, "/dev/null", std::atoi(argv[3]));
        const double t1 = now_s();
        if (t1 - t0 < best)
          best = t1 - t0;
      }
      std::printf("{\"op\": \"decompress\", \"seconds\": %.6f, \"threads\": %d}\n", best,
                  std::atoi(argv[3]));
      return 0;
    }
  } catch (const std::exception &e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return 1;
  }
  std::cerr << "gfaz_ref: bad command or argument count" << std::endl;
  return 2;
}
