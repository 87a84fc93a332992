// C entry points over the reference CpuPropagator<double> -- TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile (target `ref`) straight from the reference headers where
// they lie (-I <reference tree>); nothing of the reference is copied here.  The
// result, oracle/_ref/libcp3ref.so, validates the oracle restatement and writes the
// golden fixtures under tests/golden/.  It may also serve as the CPU baseline.
#include <cpupropagator.hpp>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

using namespace cudaprob3;

namespace {

// Exposes the protected host state of the reference propagator (propagator.hpp:317-336).
struct Probe : public CpuPropagator<double> {
    Probe(int nc, int ne, int threads) : CpuPropagator<double>(nc, ne, threads) {}
    const std::array<math::ComplexNumber<double>, 9>& mix() const { return this->Mix_U; }
    const std::array<double, 9>& masses() const { return this->dm; }
    const std::vector<int>& layers() const { return this->maxlayers; }
    const std::vector<double>& limits() const { return this->coslimit; }
    const std::vector<double>& rad() const { return this->radii; }
    const std::vector<double>& rho() const { return this->rhos; }
    double height_cm() const { return this->ProductionHeightinCentimeter; }
};

}  // namespace

extern "C" {

// Full CpuPropagator<double> run; result is AoS [(ic*n_e+ie)*9 + t] (cpupropagator.hpp:101-103).
// Returns 0, or 1 if the reference threw.
int ref_propagate(int antinu, const double* cosines, int n_c, const double* energies, int n_e,
                  const double* radii, const double* rhos, int n_rows, double h_km,
                  double t12, double t13, double t23, double dcp, double dm12sq, double dm23sq,
                  int threads, double* result) {
    try {
        Probe prop(n_c, n_e, threads > 0 ? threads : 1);
        prop.setEnergyList(std::vector<double>(energies, energies + n_e));
        prop.setCosineList(std::vector<double>(cosines, cosines + n_c));
        prop.setMNSMatrix(t12, t13, t23, dcp);
        prop.setNeutrinoMasses(dm12sq, dm23sq);
        prop.setDensity(std::vector<double>(radii, radii + n_rows), std::vector<double>(rhos, rhos + n_rows));
        prop.setProductionHeight(h_km);
        prop.calculateProbabilities(antinu ? Antineutrino : Neutrino);
        for (int ic = 0; ic < n_c; ic++)
            for (int ie = 0; ie < n_e; ie++)
                for (int t = 0; t < 9; t++)
                    result[(std::uint64_t(ic) * n_e + ie) * 9 + t] = prop.getProbability(ic, ie, ProbType(t));
    } catch (const std::exception&) {
        return 1;
    }
    return 0;
}

// Host set-up of the reference: U (re,im interleaved), DM, maxlayers, coslimit,
// descending radii/rhos and the production height in cm.
int ref_setup(const double* cosines, int n_c, const double* radii, const double* rhos, int n_rows,
              double h_km, double t12, double t13, double t23, double dcp, double dm12sq, double dm23sq,
              double* U18, double* DM9, int* maxlayers, double* coslimit, double* radii_desc,
              double* rhos_desc, double* h_cm) {
    try {
        Probe prop(n_c, 1, 1);
        prop.setCosineList(std::vector<double>(cosines, cosines + n_c));
        prop.setMNSMatrix(t12, t13, t23, dcp);
        prop.setNeutrinoMasses(dm12sq, dm23sq);
        prop.setDensity(std::vector<double>(radii, radii + n_rows), std::vector<double>(rhos, rhos + n_rows));
        prop.setProductionHeight(h_km);
        for (int i = 0; i < 9; i++) { U18[2 * i] = prop.mix()[i].re; U18[2 * i + 1] = prop.mix()[i].im; }
        for (int i = 0; i < 9; i++) DM9[i] = prop.masses()[i];
        for (int i = 0; i < n_c; i++) maxlayers[i] = prop.layers()[i];
        for (int i = 0; i < n_rows; i++) {
            coslimit[i] = prop.limits()[i];
            radii_desc[i] = prop.rad()[i];
            rhos_desc[i] = prop.rho()[i];
        }
        *h_cm = prop.height_cm();
    } catch (const std::exception&) {
        return 1;
    }
    return 0;
}

// One layer's transition matrix through the reference physics (physics.hpp:522-531).
void ref_transition_matrix(const double* U18, const double* DM9, int antinu, double E, double rho,
                           double Lkm, double* A18) {
    math::ComplexNumber<double> U[9];
    double dm[9];
    for (int i = 0; i < 9; i++) { U[i].re = U18[2 * i]; U[i].im = U18[2 * i + 1]; dm[i] = DM9[i]; }
    physics::setMixMatrix_host(U);
    physics::setMassDifferences_host(dm);
    const NeutrinoType type = antinu ? Antineutrino : Neutrino;
    physics::prepare_getMfast<double>(type);
    math::ComplexNumber<double> A[3][3];
    physics::get_transition_matrix<double>(type, E, rho, Lkm, A, 0.0);
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) { A18[2 * (n * 3 + m)] = A[n][m].re; A18[2 * (n * 3 + m) + 1] = A[n][m].im; }
}

// Reference eigenvalue routine (physics.hpp:268-341).
void ref_matter_masses(const double* U18, const double* DM9, double E, double rho, int antinu,
                       double* mm9, double* mv9) {
    math::ComplexNumber<double> U[9];
    double dm[9];
    for (int i = 0; i < 9; i++) { U[i].re = U18[2 * i]; U[i].im = U18[2 * i + 1]; dm[i] = DM9[i]; }
    physics::setMixMatrix_host(U);
    physics::setMassDifferences_host(dm);
    const NeutrinoType type = antinu ? Antineutrino : Neutrino;
    physics::prepare_getMfast<double>(type);
    double mm[3][3], mv[3][3];
    physics::getMfast<double>(E, rho, type, mm, mv);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { mm9[i * 3 + j] = mm[i][j]; mv9[i * 3 + j] = mv[i][j]; }
}

}  // extern "C"
