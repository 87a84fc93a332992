// Single-GPU propagator on an MI355X, a drop-in for the reference's
// CudaPropagatorSingle<FLOAT_T> (cudapropagator.cuh:39-251).
//
// Header-only C++ over the C ABI of libcp3hip.so (include/cp3hip.h): the base class does the
// host set-up, this class forwards the results to the device handle, launches, and copies
// the probabilities back lazily, into the handle's pinned host buffer, on the first
// getProbability after a computation, like the reference (cudapropagator.cuh:73, 172-175,
// 229-235).  Link with -lcp3hip.  Device failures throw std::runtime_error (the reference
// prints and exits).
#ifndef CUDAPROB3_MI355X_HIPPROPAGATOR_HPP
#define CUDAPROB3_MI355X_HIPPROPAGATOR_HPP

#include "../cp3hip.h"
#include "propagator.hpp"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cudaprob3 {

// Arithmetic of the device kernels.
enum class Precision {
    FP64Plain = CP3H_FP64_PLAIN,   // the reference's per-pair FP64 Barger step
    FP64 = CP3H_FP64,              // FP64, eigen-decomposition once per (shell, energy)
    Mixed = CP3H_MIXED             // FP32 bulk, FP64 eigenvalues and phases, float results
};

namespace detail {
inline void check(int status, const char* where) {
    if (status != CP3H_OK) throw std::runtime_error(std::string(where) + ": " + cp3h_last_error());
}
inline int set_energies(cp3h_propagator* h, const std::vector<double>& v) {
    return cp3h_set_energies_f64(h, v.data(), int(v.size()));
}
inline int set_energies(cp3h_propagator* h, const std::vector<float>& v) {
    return cp3h_set_energies_f32(h, v.data(), int(v.size()));
}
inline int set_cosines(cp3h_propagator* h, const std::vector<double>& v, const std::vector<int>& ml) {
    return cp3h_set_cosines_f64(h, v.data(), ml.data(), int(v.size()));
}
inline int set_cosines(cp3h_propagator* h, const std::vector<float>& v, const std::vector<int>& ml) {
    return cp3h_set_cosines_f32(h, v.data(), ml.data(), int(v.size()));
}
inline int set_density(cp3h_propagator* h, const std::vector<double>& r, const std::vector<double>& d) {
    return cp3h_set_density_f64(h, r.data(), d.data(), int(r.size()));
}
inline int set_density(cp3h_propagator* h, const std::vector<float>& r, const std::vector<float>& d) {
    return cp3h_set_density_f32(h, r.data(), d.data(), int(r.size()));
}
inline int set_params(cp3h_propagator* h, const double* U18, const double* DM9) {
    return cp3h_set_params_f64(h, U18, DM9);
}
inline int set_params(cp3h_propagator* h, const float* U18, const float* DM9) {
    return cp3h_set_params_f32(h, U18, DM9);
}
}  // namespace detail

template <class FLOAT_T>
class HipPropagatorSingle : public Propagator<FLOAT_T> {
public:
    static constexpr Precision default_precision() {
        return sizeof(FLOAT_T) == sizeof(double) ? Precision::FP64 : Precision::Mixed;
    }

    // cudapropagator.cuh:50: device id, grid size.  The precision argument is an extension.
    HipPropagatorSingle(int id, int n_cosines_, int n_energies_, Precision precision = default_precision())
        : Propagator<FLOAT_T>(n_cosines_, n_energies_), deviceId(id) {
        detail::check(cp3h_create(id, n_cosines_, n_energies_, int(precision), &handle),
                      "HipPropagatorSingle");
    }

    // cudapropagator.cuh:86: device 0.
    HipPropagatorSingle(int n_cosines_, int n_energies_) : HipPropagatorSingle(0, n_cosines_, n_energies_) {}

    ~HipPropagatorSingle() override { cp3h_destroy(handle); }

    HipPropagatorSingle(const HipPropagatorSingle&) = delete;
    HipPropagatorSingle& operator=(const HipPropagatorSingle&) = delete;

    HipPropagatorSingle(HipPropagatorSingle&& other) : Propagator<FLOAT_T>(0, 0) { *this = std::move(other); }

    HipPropagatorSingle& operator=(HipPropagatorSingle&& other) {
        if (this != &other) {
            Propagator<FLOAT_T>::operator=(std::move(other));
            cp3h_destroy(handle);
            handle = other.handle;
            other.handle = nullptr;
            hostResults = other.hostResults;
            resultBytes = other.resultBytes;
            other.hostResults = nullptr;
            deviceId = other.deviceId;
            resultsResideOnHost = other.resultsResideOnHost;
        }
        return *this;
    }

    void setDensity(const std::vector<FLOAT_T>& radii_, const std::vector<FLOAT_T>& rhos_) override {
        Propagator<FLOAT_T>::setDensity(radii_, rhos_);
        detail::check(detail::set_density(live(), this->radii, this->rhos), "HipPropagatorSingle::setDensity");
    }

    void setEnergyList(const std::vector<FLOAT_T>& list) override {
        Propagator<FLOAT_T>::setEnergyList(list);
        detail::check(detail::set_energies(live(), this->energyList), "HipPropagatorSingle::setEnergyList");
    }

    void setCosineList(const std::vector<FLOAT_T>& list) override {
        Propagator<FLOAT_T>::setCosineList(list);
        detail::check(detail::set_cosines(live(), this->cosineList, this->maxlayers),
                      "HipPropagatorSingle::setCosineList");
    }

    // cudapropagator.cuh:162-165
    void calculateProbabilities(NeutrinoType type) override {
        calculateProbabilitiesAsync(type);
        waitForCompletion();
    }

    // cudapropagator.cuh:168-181; results are SoA on the device and on the host.
    FLOAT_T getProbability(int index_cosine, int index_energy, ProbType t) override {
        if (index_cosine >= this->n_cosines || index_energy >= this->n_energies || index_cosine < 0 ||
            index_energy < 0)
            throw std::runtime_error("HipPropagatorSingle::getProbability. Invalid indices");
        if (!resultsResideOnHost) {
            getResultFromDevice();
            resultsResideOnHost = true;
        }
        const std::uint64_t index = std::uint64_t(index_cosine) * std::uint64_t(this->n_energies) + index_energy;
        const std::uint64_t offset = std::uint64_t(t) * std::uint64_t(this->n_energies) * this->n_cosines;
        return resultBytes == 8 ? FLOAT_T(static_cast<const double*>(hostResults)[index + offset])
                                : FLOAT_T(static_cast<const float*>(hostResults)[index + offset]);
    }

    // Extensions: device-resident results (SoA, element size 4 for Precision::Mixed, else 8)
    // for downstream kernels, and the handle's stream.
    const void* deviceResults(int* elem_bytes = nullptr) {
        const void* p = nullptr;
        int b = 0;
        detail::check(cp3h_device_results(live(), &p, &b), "HipPropagatorSingle::deviceResults");
        if (elem_bytes) *elem_bytes = b;
        return p;
    }
    void* stream() {
        void* s = nullptr;
        detail::check(cp3h_stream(live(), &s), "HipPropagatorSingle::stream");
        return s;
    }

    // cudapropagator.cuh:191-220: launch without waiting.
    void calculateProbabilitiesAsync(NeutrinoType type) {
        if (!this->isInit || !handle)
            throw std::runtime_error("HipPropagatorSingle::calculateProbabilities. Object has been moved from.");
        if (!this->isSetProductionHeight)
            throw std::runtime_error("HipPropagatorSingle::calculateProbabilities. production height was not set");
        resultsResideOnHost = false;
        FLOAT_T U18[18];
        for (int i = 0; i < 9; i++) {
            U18[2 * i] = this->Mix_U[i].re;
            U18[2 * i + 1] = this->Mix_U[i].im;
        }
        detail::check(detail::set_params(handle, U18, this->dm.data()), "HipPropagatorSingle::calculateProbabilities");
        detail::check(cp3h_compute_async(handle, type == Antineutrino ? 1 : 0,
                                         double(this->ProductionHeightinCentimeter)),
                      "HipPropagatorSingle::calculateProbabilities");
    }

    // cudapropagator.cuh:223-226
    void waitForCompletion() { detail::check(cp3h_wait(live()), "HipPropagatorSingle::waitForCompletion"); }

protected:
    void setMaxlayers() override {
        Propagator<FLOAT_T>::setMaxlayers();
        if (handle)
            detail::check(cp3h_set_maxlayers(handle, this->maxlayers.data(), this->n_cosines),
                          "HipPropagatorSingle::setMaxlayers");
    }

    // cudapropagator.cuh:229-235: D2H into the handle's pinned host buffer
    void getResultFromDevice() {
        detail::check(cp3h_fetch_results(live(), &hostResults, &resultBytes), "HipPropagatorSingle::getProbability");
    }

private:
    cp3h_propagator* live() {
        if (!handle) throw std::runtime_error("HipPropagatorSingle: object has been moved from");
        return handle;
    }

    cp3h_propagator* handle = nullptr;
    const void* hostResults = nullptr;   // pinned, owned by the handle; SoA [t][ic][ie]
    int resultBytes = 8;                 // 8 (FP64 modes) or 4 (Precision::Mixed)
    int deviceId = 0;
    bool resultsResideOnHost = false;
};

}  // namespace cudaprob3

#endif
