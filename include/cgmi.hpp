// cgmi.hpp -- the reference's C++ driver surface over the C ABI of
// libcgmi.so (header only).  Class and member names follow the reference
// (paths relative to CGgraph/):
//   CSR_Result_type           Basic/Graph/basic_struct.hpp:7-16
//   Algorithm_type, Engine_type, OrderMethod   basic_struct.hpp:41-178
//   TwoLevelPartition / getPartitionResult     my_project/CGgraph/Two-level_partition.cuh:106, 725
//   CGgraphEngine / CG_co_execution / vertexValue   my_project/CGgraph/CGgraph_engine.cuh:154, 265, 69
// Errors throw cgmi::Error (status code + message) instead of exit(1).
#ifndef CGMI_HPP
#define CGMI_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cgmi.h"

namespace cgmi {

enum Algorithm_type { BFS = CG_BFS, SSSP = CG_SSSP, CC = CG_CC, PR = CG_PR, UNDEFINE = 4 };
enum Engine_type {
  COORDINATION = CG_COORDINATION,
  COOPERATION = 1,
  SINGLE_CORE = 2,
  MULTI_CORE = CG_MULTI_CORE,
  SINGLE_GPU = CG_SINGLE_GPU
};
enum OrderMethod { NATIVE = 0, RCM_LAST_ZERO_OUTDEGREE = 6 };

constexpr uint32_t VertexValue_MAX = CG_UNREACHED;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error("cgmi error " + std::to_string(c) + ": " + m), code(c) {}
};

inline void check(int status) {
  if (status != CG_OK) throw Error(status, cg_last_error());
}

// The reference's CSR container; the arrays are borrowed (never freed here).
struct CSR_Result_type {
  uint32_t vertexNum = 0;
  uint64_t edgeNum = 0;
  const uint32_t* csr_offset = nullptr;  // u32[V+1] (reference files) ...
  const uint64_t* csr_offset64 = nullptr;  // ... or u64[V+1]
  const uint32_t* csr_dest = nullptr;
  const uint32_t* csr_weight = nullptr;
  uint32_t noZeroOutDegreeNum = 0;
};

class TwoLevelPartition {
 public:
  struct Partition_type {
    uint64_t segmentNum = 0;
    bool canHoldAllSegment = true;
    uint64_t bucket_start[5] = {0, 0, 0, 0, 0};
    std::shared_ptr<cg_graph> graph;  // the graph in HBM, shared with the engine
  };

  TwoLevelPartition(const CSR_Result_type& csr, uint32_t zeroOutDegreeNum, size_t deviceNum,
                    Algorithm_type algorithm) {
    if (deviceNum != 1) throw Error(CG_ERR_UNSUPPORTED, "libcgmi drives one GPU per engine");
    cg_graph* g = nullptr;
    const void* off = csr.csr_offset64 ? static_cast<const void*>(csr.csr_offset64) : csr.csr_offset;
    check(cg_graph_from_csr(csr.vertexNum, csr.edgeNum, off, csr.csr_offset64 != nullptr, csr.csr_dest,
                            algorithm == SSSP ? csr.csr_weight : nullptr, zeroOutDegreeNum, &g));
    part_.graph.reset(g, cg_graph_destroy);
    cg_partition_info info{};
    check(cg_partition(g, algorithm, &info));
    part_.segmentNum = info.segment_num;
    part_.canHoldAllSegment = info.can_hold_all != 0;
    for (int i = 0; i < 5; ++i) part_.bucket_start[i] = info.bucket_start[i];
  }
  Partition_type getPartitionResult() const { return part_; }

 private:
  Partition_type part_;
};

class CGgraphEngine {
 public:
  CGgraphEngine(const CSR_Result_type& csr, const TwoLevelPartition::Partition_type& partition,
                Algorithm_type algorithm, int useDeviceNum, const cg_options* opt = nullptr,
                Engine_type engine_type = COORDINATION)
      : vertexNum(csr.vertexNum), graph_(partition.graph), algorithm_(algorithm) {
    cg_engine* e = nullptr;
    check(cg_engine_create(graph_.get(), algorithm, engine_type, useDeviceNum, opt, &e));
    engine_.reset(e, cg_engine_destroy);
    vertexValue.resize(vertexNum);
  }

  // Runs one query from `root`; returns the wall time in ms and leaves the
  // result in vertexValue (uint32; PageRank ranks are in vertexValue_pr).
  double CG_co_execution(uint32_t root = 0) {
    cg_run_stats st{};
    if (algorithm_ == PR) {
      vertexValue_pr.resize(vertexNum);
      check(cg_engine_run(engine_.get(), root, vertexValue_pr.data(), &st));
    } else {
      check(cg_engine_run(engine_.get(), root, vertexValue.data(), &st));
    }
    last_stats = st;
    return st.ms;
  }

  uint32_t vertexNum;
  std::vector<uint32_t> vertexValue;
  std::vector<float> vertexValue_pr;
  cg_run_stats last_stats{};

 private:
  std::shared_ptr<cg_graph> graph_;
  std::shared_ptr<cg_engine> engine_;
  Algorithm_type algorithm_;
};

}  // namespace cgmi

#endif
