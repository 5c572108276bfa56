// Command-line driver with the reference's modes and arguments (inv.cu:6-82):
//   inv.out [config_dir] [-key=value ...]
//   mode 0  inversion (misfit + optimizer)
//   mode 1  forward of every source at once, traces + wavefield snapshots
//   mode 2  misfit kernel (forward + adjoint of each shot), kernels exported
//   mode 3  trace generation, one forward per shot
//   mode 4  device query
//   mode 5  model export on the grid
// Keys beyond the reference's: batch (shots per launch, default 16), math
// (0 exact, 1 fast), graphs (0/1), device, snap_format (0 FP32, 1 FP16,
// 2 BF16, 3 I16), ckpt (checkpoint interval), autotune (0/1), lbfgs_fix.
// Mode 2 also prints the gradient's wall time and misfit on stderr.
#include <hip/hip_runtime.h>

#include <chrono>

#include "inv.h"

using namespace invhost;

namespace {

size_t getMemoryUsage() {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
    return free_b;
}

// checkMemoryUsage (lib/query.h:9-22)
void checkMemoryUsage(size_t spare_byte) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return;
    const float MB = 1024.0f * 1024.0f;
    const float free_db = (float)free_b / MB, total_db = (float)spare_byte / MB;
    std::cout.setf(std::ios_base::fixed, std::ios_base::floatfield);
    std::cout.precision(1);
    std::cout << "Memory usage: " << total_db - free_db << "MB / " << total_db << "MB" << std::endl;
}

// deviceQuery (lib/query.h:23-66), HIP device properties
void deviceQuery() {
    int n = 0;
    hipError_t err = hipGetDeviceCount(&n);
    if (err != hipSuccess) {
        std::cout << "Failed to get device info: " << hipGetErrorString(err) << std::endl;
        return;
    }
    if (n == 0) std::cout << "No HIP device available" << std::endl;
    else std::cout << n << " HIP device(s) detected" << std::endl;
    std::cout.setf(std::ios_base::fixed, std::ios_base::floatfield);
    std::cout.precision(1);
    for (int d = 0; d < n; d++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
        int drv = 0, rt = 0;
        (void)hipDriverGetVersion(&drv);
        (void)hipRuntimeGetVersion(&rt);
        std::cout << "\nDevice " << d << ": " << p.name << " (" << p.gcnArchName << ")" << std::endl;
        std::cout << "  HIP Driver Version / Runtime Version           " << drv << " / " << rt << std::endl;
        std::cout << "  Compute units / XCD-visible L2 (MB):           " << p.multiProcessorCount
                  << " / " << (float)p.l2CacheSize / 1048576.0f << std::endl;
        std::cout << "  Total amount of global memory:                 "
                  << (float)p.totalGlobalMem / 1048576.0f << " MB" << std::endl;
        std::cout << "  LDS per workgroup:                             "
                  << (float)p.sharedMemPerBlock / 1024.0f << " KB" << std::endl;
        std::cout << "  Maximum number of threads per block:           " << p.maxThreadsPerBlock << std::endl;
        std::cout << "  Maximum sizes of each dimension of a block:    " << p.maxThreadsDim[0] << " x "
                  << p.maxThreadsDim[1] << " x " << p.maxThreadsDim[2] << std::endl;
        std::cout << "  Maximum sizes of each dimension of a grid:     " << p.maxGridSize[0] << " x "
                  << p.maxGridSize[1] << " x " << p.maxGridSize[2] << std::endl;
    }
}

template <class T>
std::unique_ptr<T> need(T *p, const char *what) {
    if (!p) throw std::runtime_error(string("unknown ") + what + " id");
    return std::unique_ptr<T>(p);
}

int run(int argc, const char *argv[]) {
    map<string, string> cfg;
    for (int a = 1; a < argc; a++) {
        const string arg = argv[a];
        const size_t pos = arg.find('=');
        if (arg[0] == '-') {
            if (pos != string::npos) cfg[arg.substr(1, pos - 1)] = arg.substr(pos + 1);
        } else {
            cfg["config"] = arg;
        }
    }
    if (!cfg["config"].size()) {
        std::cout << "Using default directory example/checker" << std::endl;
        cfg["config"] = "example/checker";
    }
    Config config(cfg);
    const int mode = config.i["mode"];
    if (mode == 4) {
        deviceQuery();
        return 0;
    }
    const size_t free_memory = getMemoryUsage();
    switch (mode) {
        case 0: {
            auto solver = need(module::solver(config.i["solver"]), "solver");
            std::unique_ptr<Filter> filter(module::filter(config.i["filter"]));
            auto misfit = need(module::misfit(config.i["misfit"]), "misfit");
            auto optimizer = need(module::optimizer(config.i["optimizer"]), "optimizer");
            misfit->init(&config, solver.get(), filter.get());
            optimizer->init(&config, solver.get(), misfit.get());
            optimizer->run();
            checkMemoryUsage(free_memory);
            break;
        }
        case 1: {
            auto solver = need(module::solver(config.i["solver"]), "solver");
            solver->init(&config);
            solver->importModel(true);
            solver->exportAxis();
            solver->runForward(-1, true, true);
            checkMemoryUsage(free_memory);
            break;
        }
        case 2: {
            auto solver = need(module::solver(config.i["solver"]), "solver");
            std::unique_ptr<Filter> filter(module::filter(config.i["filter"]));
            auto misfit = need(module::misfit(config.i["misfit"]), "misfit");
            misfit->init(&config, solver.get(), filter.get());
            // the gradient's wall time on stderr (the reference's mode 2
            // prints nothing; stdout stays as it is)
            const auto t0 = std::chrono::steady_clock::now();
            const double f = misfit->run(true, true);
            if (auto *fdm = dynamic_cast<FdmSolver *>(solver.get())) INVH(fdm->ctx, inv_sync(fdm->ctx));
            const double secs =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::cerr << "gradient: " << secs << " s, misfit " << f << ", " << solver->nsrc
                      << " shots" << std::endl;
            solver->exportAxis();
            solver->exportKernel();
            checkMemoryUsage(free_memory);
            break;
        }
        case 3: {
            auto solver = need(module::solver(config.i["solver"]), "solver");
            solver->init(&config);
            solver->importModel(true);
            for (size_t is = 0; is < solver->nsrc; is++) solver->runForward((int)is, true);
            break;
        }
        case 5: {
            auto solver = need(module::solver(config.i["solver"]), "solver");
            solver->init(&config);
            solver->importModel(true);
            solver->exportAxis();
            solver->exportModel();
            break;
        }
        default: break;
    }
    return 0;
}

}  // namespace

int main(int argc, const char *argv[]) {
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
}
