// C++ driver in the shape of the reference's main_partition
// (my_project/CGgraph/main_partition.hpp:162-243): toy graph, partition,
// engine, `runs` executions with the trimmed mean, then the values.  The
// pytest wrapper (tests/test_gpu_cpp.py) compares the printed values with the
// golden vectors.  Usage: engine_test <toy_0|toy_1> <BFS|SSSP> <root> <runs>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/cgmi.hpp"

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  const bool toy1 = std::strcmp(argv[1], "toy_1") == 0;
  // Basic/Graph/graphFileList.hpp:510-570
  const std::vector<uint32_t> off0 = {0, 3, 6, 6, 8, 8, 11, 15, 16, 21, 22, 24, 26};
  const std::vector<uint32_t> dst0 = {1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 5, 0, 2, 4, 9, 11, 5, 2, 8, 2, 5};
  const std::vector<uint32_t> off1 = {0, 3, 6, 6, 8, 8, 11, 16, 17, 21, 24, 26, 28};
  const std::vector<uint32_t> dst1 = {1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 10, 5, 0, 4, 9, 11, 5, 6, 8, 2, 8, 2, 5};
  const std::vector<uint32_t> w1 = {1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 5, 0, 2, 4, 9, 11, 5, 2, 8, 2, 5, 2, 5};
  const auto& off = toy1 ? off1 : off0;
  const auto& dst = toy1 ? dst1 : dst0;
  const auto& w = toy1 ? w1 : dst0;  // toy_0's weights equal its destinations
  cgmi::CSR_Result_type csr;
  csr.vertexNum = 12;
  csr.edgeNum = dst.size();
  csr.csr_offset = off.data();
  csr.csr_dest = dst.data();
  csr.csr_weight = w.data();
  csr.noZeroOutDegreeNum = 12;
  const cgmi::Algorithm_type alg = std::strcmp(argv[2], "SSSP") == 0 ? cgmi::SSSP : cgmi::BFS;
  const uint32_t root = (uint32_t)std::atoi(argv[3]);
  const int runs = std::atoi(argv[4]);
  try {
    cgmi::TwoLevelPartition part(csr, 0, 1, alg);
    cgmi::CGgraphEngine engine(csr, part.getPartitionResult(), alg, 1);
    std::vector<double> t(runs);
    for (int r = 0; r < runs; ++r) t[r] = engine.CG_co_execution(root);
    std::sort(t.begin(), t.end());
    double sum = 0;
    for (int r = runs > 2 ? 1 : 0; r < (runs > 2 ? runs - 1 : runs); ++r) sum += t[r];
    std::printf("[Total Run Time]: %f (ms)\n", sum / (runs > 2 ? runs - 2 : runs));
    for (uint32_t v = 0; v < csr.vertexNum; ++v) std::printf("%u%c", engine.vertexValue[v], v + 1 < csr.vertexNum ? ' ' : '\n');
  } catch (const cgmi::Error& e) {
    std::printf("error %d %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
