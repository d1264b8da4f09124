// loader_probe.cpp -- drives the host's C++ plugin loader
// (quasi_amd/host/loader.hpp) on a plugin given on the command line and
// prints the outcome: "ok <abi> <name>" or "error <dynamic_library|loader>
// <message>".  Built and run by tests/test_host_io.py (CPU only).
#include <cstdio>

#include "loader.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    auto lib = Q::plugin::dynamic_library::open(argv[1]);
    if (!lib) {
        std::printf("error dynamic_library %s\n", std::string(Q::plugin::to_string(lib.error())).c_str());
        return 0;
    }
    Q_gpu_context gpu{Q_GPU_BACKEND_HIP, nullptr, nullptr, nullptr};
    Q_plugin_context ctx{64, 64, nullptr, &gpu, nullptr, nullptr};
    auto p = Q::plugin::loader::load(*lib, &ctx);
    if (!p) {
        std::printf("error loader %s\n", Q::plugin::to_string(p.error()));
        return 0;
    }
    std::printf("ok %u %s readback=%d aov=%d\n", p->abi_version(), p->info().name, int(p->supports_readback()),
                int(p->supports_readback_aov()));
    p->destroy();
    return 0;
}
