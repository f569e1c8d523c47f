// om_forward: simulated sensor data (apps/forward.cpp, Forward forward.h:
// 14-36): GainMatrix x RealSourcesData plus Gaussian noise of standard
// deviation NoiseLevel, the product on the GPU.
//   om_forward GainMatrix RealSourcesData SimulatedData NoiseLevel
#include <cstdlib>
#include <iostream>
#include <random>

#include "cli.hpp"

using namespace cli;

int main(int argc, char** argv) {
    if (argc < 2 || std::string(argv[1]) == "-h" || std::string(argv[1]) == "--help") {
        std::cout << argv[0] << " GainMatrix RealSourcesData SimulatedData NoiseLevel\n";
        return 0;
    }
    if (argc < 5) {
        std::cerr << "Not enough arguments." << std::endl;
        return 1;
    }
    try {
        const Full G = read_full(argv[1]);
        const Full X = read_full(argv[2]);
        const double noise = std::atof(argv[4]);
        if (G.ncol != X.nlin) throw std::runtime_error("the gain's columns do not match the sources' rows");
        Dev<double> dG(G.a.data(), G.a.size()), dX(X.a.data(), X.a.size());
        Dev<double> dY((size_t)(G.nlin * X.ncol));
        check(omh_gemm_f64(0, 0, G.nlin, X.ncol, G.ncol, 1.0, dG.p, G.nlin, dX.p, X.nlin, 0.0, dY.p, G.nlin, nullptr));
        Full Y;
        Y.nlin = G.nlin;
        Y.ncol = X.ncol;
        Y.a = dY.down();
        if (noise > 0) {
            std::random_device rd{};
            std::mt19937 gen{rd()};
            std::normal_distribution<> nd{0, noise};
            for (double& y : Y.a) y += nd(gen);
        }
        write_full(argv[3], Y);
    } catch (const std::exception& e) {
        std::cerr << argv[0] << ": " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
