// Device-driven BFS level loops (filled in below).
#include "bfs_engine.hpp"

namespace ggpu {

void BfsEngine::device_loop(int variant, uint32_t src, const ggpu_bfs_opts& o, std::vector<ggpu_level>* trace) {
    (void)variant; (void)src; (void)o; (void)trace;
    throw ArgError("device-driven loops are not built yet");
}

}  // namespace ggpu
