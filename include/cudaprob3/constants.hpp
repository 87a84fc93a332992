// Physical constants of the propagator API (the reference's constants.hpp:25-41; the
// values are the reference's and must stay identical for parity).
#ifndef CUDAPROB3_MI355X_CONSTANTS_HPP
#define CUDAPROB3_MI355X_CONSTANTS_HPP

namespace cudaprob3 {

template <typename FLOAT_T>
struct Constants {
    // 2 sqrt(2) G_F N_A in eV^2 per (GeV g/cm^3): matter potential scale
    static constexpr FLOAT_T tworttwoGf() { return 1.52588e-4; }
    static constexpr FLOAT_T km2cm() { return 1.0e5; }
    static constexpr FLOAT_T REarth() { return 6371.0; }
    static constexpr FLOAT_T REarthcm() { return REarth() * km2cm(); }
    // electrons per nucleon
    static constexpr FLOAT_T density_convert() { return 0.5; }
};

}  // namespace cudaprob3

#endif
