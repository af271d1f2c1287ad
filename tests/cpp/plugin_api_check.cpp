// Exercises include/cusuperhelium.hpp the way a C++ caller of the reference's plugin API does
// (WaterBoundaryProblem + BaseBoundaryIntegralCalculator + AutonomousRungeKuttaStepper +
// TrajectoryLogger), and writes what it computed for tests/test_gpu_cpp_api.py to check.
//
//   plugin_api_check <in.bin> <out.bin>
// in.bin : 2 N B complex128, the initial state [Z(N B), Phi(N B) + 0i] of B = 2 surfaces, N = 256
// out.bin: rhs of the initial state through run() (2 N B complex), the 10 logged states of
//          runEvolution(0, 10 dt) from a host state (10 x 2 N B complex) and their times (10 doubles),
//          and the state after 10 runStep calls on a device state (2 N B complex)
#include <cstdio>
#include <vector>

#include <hip/hip_runtime.h>

#include "cusuperhelium.hpp"

using namespace cusuperhelium;

constexpr int N = 256;
constexpr size_t B = 2;
constexpr int M = 2 * N * (int)B;

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    std::vector<std_complex> y0(M);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(y0.data(), sizeof(std_complex), M, f) != (size_t)M) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::fclose(f);
    try {
        ProblemProperties props;
        props.rho = 0.0;
        WaterBoundaryProblem<N, B> problem(props);
        BaseBoundaryIntegralCalculator<N, B> calc(props, problem);
        const double dt = 0.01;

        // run(): one right-hand side on device arrays
        std_complex *dstate = nullptr, *drhs = nullptr;
        if (hipMalloc(&dstate, M * sizeof(std_complex)) != hipSuccess ||
            hipMalloc(&drhs, M * sizeof(std_complex)) != hipSuccess)
            throw std::runtime_error("hipMalloc failed");
        if (hipMemcpy(dstate, y0.data(), M * sizeof(std_complex), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        calc.run(dstate, drhs);
        std::vector<std_complex> rhs(M);
        if (hipMemcpy(rhs.data(), drhs, M * sizeof(std_complex), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");

        // runEvolution from a host state with a trajectory logger
        auto logger = std::make_shared<TrajectoryLogger<std_complex, M>>();
        AutonomousRungeKuttaStepper<std_complex, M> stepper(calc, dt, logger);
        stepper.initialize(y0.data(), false);
        stepper.runEvolution(0.0, 10 * dt + 1e-12);
        std_complex* states = nullptr;
        double* times = nullptr;
        size_t ns = 0, nt = 0;
        if (logger->copyStatesToHost(&states, &ns) != 0 || logger->copyTimesToHost(&times, &nt) != 0)
            throw std::runtime_error("logger copy failed");
        if (ns != 10 || nt != 10) throw std::runtime_error("expected 10 logged states");

        // runStep on a device state, advanced in place
        AutonomousRungeKuttaStepper<std_complex, M> stepper2(calc);
        RK4Options opt;
        opt.initial_timestep = dt;
        stepper2.setOptions(opt);
        stepper2.initialize(dstate, true);
        for (int i = 0; i < 10; ++i) stepper2.runStep(i);
        std::vector<std_complex> yd(M);
        if (hipMemcpy(yd.data(), dstate, M * sizeof(std_complex), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        (void)hipFree(dstate);
        (void)hipFree(drhs);

        FILE* o = std::fopen(argv[2], "wb");
        if (!o) throw std::runtime_error("cannot write output");
        std::fwrite(rhs.data(), sizeof(std_complex), M, o);
        std::fwrite(states, sizeof(std_complex), (size_t)M * ns, o);
        std::fwrite(times, sizeof(double), nt, o);
        std::fwrite(yd.data(), sizeof(std_complex), M, o);
        std::fclose(o);
        std::free(states);
        std::free(times);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "plugin_api_check: %s\n", e.what());
        return 1;
    }
    std::printf("plugin_api_check: ok\n");
    return 0;
}
