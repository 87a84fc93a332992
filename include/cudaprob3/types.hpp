// Enumerations of the propagator API (same names and values as the reference's
// types.hpp:23-29, so code written against it compiles unchanged).
#ifndef CUDAPROB3_MI355X_TYPES_HPP
#define CUDAPROB3_MI355X_TYPES_HPP

namespace cudaprob3 {

// P(in -> out) is addressed by in*3 + out.
enum ProbType : int {
    e_e = 0, e_m = 1, e_t = 2,
    m_e = 3, m_m = 4, m_t = 5,
    t_e = 6, t_m = 7, t_t = 8
};

// Antineutrino flips the sign of the matter potential (and nothing else).
enum NeutrinoType { Neutrino, Antineutrino };

}  // namespace cudaprob3

#endif
