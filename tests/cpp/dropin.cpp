// Drop-in check of the C++ API: the call sequence of the reference example
// (example/main.cpp:107-139) against HipPropagatorSingle<FLOAT_T> through a
// std::unique_ptr<Propagator<FLOAT_T>>, with only the class name swapped.
//
// usage: dropin <spec file>
//   spec: n_cosines n_energies antinu h_km type(f64|f32) precision(0|1|2)
//         theta12 theta13 theta23 dcp dm12sq dm23sq
//         <n_cosines cosines> <n_energies energies>
//         <density file> <output file>
// Writes P[ic][ie][t] as float64 (getProbability, t = in*3+out) to the output file.
#include <cudaprob3/hippropagator.hpp>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

using namespace cudaprob3;

template <class T>
int run(std::ifstream& spec, int nc, int ne, int antinu, double h_km, int precision) {
    double p[6];
    for (double& x : p) spec >> x;
    std::vector<T> cosines(nc), energies(ne);
    for (auto& c : cosines) { double v; spec >> v; c = T(v); }
    for (auto& e : energies) { double v; spec >> v; e = T(v); }
    std::string density, output;
    spec >> density >> output;

    std::unique_ptr<Propagator<T>> propagator(
        new HipPropagatorSingle<T>(0, nc, ne, static_cast<Precision>(precision)));
    propagator->setEnergyList(energies);
    propagator->setCosineList(cosines);
    propagator->setMNSMatrix(T(p[0]), T(p[1]), T(p[2]), T(p[3]));
    propagator->setNeutrinoMasses(T(p[4]), T(p[5]));
    propagator->setDensityFromFile(density);
    propagator->setProductionHeight(T(h_km));
    propagator->calculateProbabilities(antinu ? Antineutrino : Neutrino);

    std::vector<double> out(std::size_t(nc) * ne * 9);
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < ne; j++)
            for (int t = 0; t < 9; t++)
                out[(std::size_t(i) * ne + j) * 9 + t] = propagator->getProbability(i, j, ProbType(t));
    std::FILE* f = std::fopen(output.c_str(), "wb");
    if (!f) return 3;
    std::fwrite(out.data(), sizeof(double), out.size(), f);
    std::fclose(f);

    // the API's error conventions
    try {
        propagator->getProbability(nc, 0, e_e);
        return 4;
    } catch (const std::runtime_error&) {
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: dropin <spec>\n";
        return 2;
    }
    std::ifstream spec(argv[1]);
    int nc, ne, antinu, precision;
    double h_km;
    std::string type;
    spec >> nc >> ne >> antinu >> h_km >> type >> precision;
    try {
        return type == "f32" ? run<float>(spec, nc, ne, antinu, h_km, precision)
                             : run<double>(spec, nc, ne, antinu, h_km, precision);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
