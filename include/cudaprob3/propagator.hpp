// Abstract propagator: host-side set-up shared by every backend.
//
// Same public interface as the reference's cudaprob3::Propagator<FLOAT_T>
// (propagator.hpp:43-337): setters for the grid, the mixing matrix, the mass splittings,
// the Earth model and the production height, and the two pure virtuals
// calculateProbabilities / getProbability.  The set-up arithmetic (PMNS matrix, mass
// differences with the 5e-9 degeneracy nudge, density flip, coslimit, maxlayers) uses the
// reference's formulas in the same order and precision, because the device results depend
// on these bits; the error messages are the reference's too.
#ifndef CUDAPROB3_MI355X_PROPAGATOR_HPP
#define CUDAPROB3_MI355X_PROPAGATOR_HPP

#include "constants.hpp"
#include "types.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cudaprob3 {

namespace math {
template <typename T>
struct ComplexNumber {
    T re;
    T im;
};
}  // namespace math

template <class FLOAT_T>
class Propagator {
public:
    Propagator(int n_cosines_, int n_energies_) : n_cosines(n_cosines_), n_energies(n_energies_) {
        energyList.resize(n_energies);
        cosineList.resize(n_cosines);
        maxlayers.resize(n_cosines);
    }

    Propagator(const Propagator& other) { *this = other; }
    Propagator(Propagator&& other) { *this = std::move(other); }

    Propagator& operator=(const Propagator& other) {
        copy_state(other);
        isInit = other.isInit;
        return *this;
    }

    Propagator& operator=(Propagator&& other) {
        copy_state(other);
        isInit = other.isInit;
        other.isInit = false;   // a moved-from propagator refuses to compute
        return *this;
    }

    virtual ~Propagator() = default;

    // Density (g/cm^3) between radii[i-1] (exclusive) and radii[i] (inclusive), radii in km,
    // both lists in the same order, ascending or descending.  propagator.hpp:117-161.
    virtual void setDensity(const std::vector<FLOAT_T>& radii_, const std::vector<FLOAT_T>& rhos_) {
        if (rhos_.size() != radii_.size()) throw std::runtime_error("setDensity : rhos.size() != radii.size()");
        if (rhos_.empty()) throw std::runtime_error("setDensity : vectors must not be empty");
        bool ascending = false;
        if (radii_.size() >= 2) {
            const int sign = radii_[1] - radii_[0] > 0 ? 1 : -1;
            for (std::size_t i = 1; i < radii_.size(); i++)
                if ((radii_[i] - radii_[i - 1]) * sign < 0) throw std::runtime_error("radii order messed up");
            ascending = sign == 1;
        }
        radii = radii_;
        rhos = rhos_;
        if (ascending) {     // keep the outermost shell first
            std::reverse(radii.begin(), radii.end());
            std::reverse(rhos.begin(), rhos.end());
        }
        coslimit.assign(radii.size(), FLOAT_T(0));
        const FLOAT_T R = Constants<FLOAT_T>::REarth();
        for (std::size_t i = 1; i < radii.size(); i++) {
            // cosine below which a path enters the shell of outer radius radii[i]
            const FLOAT_T q = 1 - (radii[i] * radii[i] / (R * R));
            coslimit[i] = FLOAT_T(-1 * std::sqrt(double(q)));
        }
        setMaxlayers();
    }

    // Two columns: radius (km, starting at 0) and density (g/cm^3).  propagator.hpp:169-184.
    virtual void setDensityFromFile(const std::string& filename) {
        std::ifstream file(filename);
        if (!file) throw std::runtime_error("could not open density file " + filename);
        std::vector<FLOAT_T> r_list, d_list;
        FLOAT_T r, d;
        while (file >> r >> d) {
            r_list.push_back(r);
            d_list.push_back(d);
        }
        setDensity(r_list, d_list);
    }

    // Mixing angles and CP phase in radians, PDG parametrisation.  propagator.hpp:191-221.
    virtual void setMNSMatrix(FLOAT_T theta12, FLOAT_T theta13, FLOAT_T theta23, FLOAT_T dCP) {
        const FLOAT_T s12 = sind(theta12), s13 = sind(theta13), s23 = sind(theta23);
        const FLOAT_T c12 = cosd(theta12), c13 = cosd(theta13), c23 = cosd(theta23);
        const FLOAT_T sd = sind(dCP), cd = cosd(dCP);
        set_u(0, 0, c12 * c13, 0.0);
        set_u(0, 1, s12 * c13, 0.0);
        set_u(0, 2, s13 * cd, -s13 * sd);
        set_u(1, 0, -s12 * c23 - c12 * s23 * s13 * cd, -c12 * s23 * s13 * sd);
        set_u(1, 1, c12 * c23 - s12 * s23 * s13 * cd, -s12 * s23 * s13 * sd);
        set_u(1, 2, s23 * c13, 0.0);
        set_u(2, 0, s12 * s23 - c12 * c23 * s13 * cd, -c12 * c23 * s13 * sd);
        set_u(2, 1, -c12 * s23 - s12 * c23 * s13 * cd, -s12 * c23 * s13 * sd);
        set_u(2, 2, c23 * c13, 0.0);
    }

    // Mass-squared splittings in eV^2, no ordering assumed.  propagator.hpp:226-248.
    virtual void setNeutrinoMasses(FLOAT_T dm12sq, FLOAT_T dm23sq) {
        FLOAT_T m[3] = {FLOAT_T(0.0), dm12sq, dm12sq + dm23sq};
        const FLOAT_T nudge = 5.0e-9;     // separates exactly degenerate masses
        if (dm12sq == 0.0) m[0] -= nudge;
        if (dm23sq == 0.0) m[2] += nudge;
        for (int i = 0; i < 3; i++) DM(i, i) = 0.0;
        DM(0, 1) = m[0] - m[1];
        DM(1, 0) = -DM(0, 1);
        DM(0, 2) = m[0] - m[2];
        DM(2, 0) = -DM(0, 2);
        DM(1, 2) = m[1] - m[2];
        DM(2, 1) = -DM(1, 2);
    }

    // Energies in GeV.  propagator.hpp:252-257.
    virtual void setEnergyList(const std::vector<FLOAT_T>& list) {
        if (list.size() != std::size_t(n_energies))
            throw std::runtime_error(
                "Propagator::setEnergyList. Propagator was not created for this number of energy nodes");
        energyList = list;
    }

    // Cosines of the zenith angle.  propagator.hpp:261-274.
    virtual void setCosineList(const std::vector<FLOAT_T>& list) {
        if (list.size() != std::size_t(n_cosines))
            throw std::runtime_error(
                "Propagator::setCosineList. Propagator was not created for this number of cosine nodes");
        cosineList = list;
        if (isSetProductionHeight) setProductionHeight(ProductionHeightinCentimeter / 100000.0);
        setMaxlayers();
        isSetCosine = true;
    }

    // Height of the production point above the surface, km (a vacuum layer).  propagator.hpp:279-286.
    virtual void setProductionHeight(FLOAT_T heightKM) {
        if (!isSetCosine) throw std::runtime_error("must set cosine list before production height");
        ProductionHeightinCentimeter = heightKM * 100000.0;
        isSetProductionHeight = true;
    }

    virtual void calculateProbabilities(NeutrinoType type) = 0;

    virtual FLOAT_T getProbability(int index_cosine, int index_energy, ProbType t) = 0;

protected:
    // Shells crossed at each cosine: #{i : cos < coslimit[i]}.  propagator.hpp:301-307.
    virtual void setMaxlayers() {
        for (int ic = 0; ic < n_cosines; ic++) {
            const FLOAT_T c = cosineList[ic];
            maxlayers[ic] = int(std::count_if(coslimit.begin(), coslimit.end(),
                                              [c](FLOAT_T limit) { return c < limit; }));
        }
    }

    math::ComplexNumber<FLOAT_T>& U(int i, int j) { return Mix_U[i * 3 + j]; }
    FLOAT_T& DM(int i, int j) { return dm[i * 3 + j]; }

    std::vector<FLOAT_T> energyList;
    std::vector<FLOAT_T> cosineList;
    std::vector<int> maxlayers;
    std::vector<FLOAT_T> radii;
    std::vector<FLOAT_T> rhos;
    std::vector<FLOAT_T> coslimit;
    std::array<math::ComplexNumber<FLOAT_T>, 9> Mix_U{};
    std::array<FLOAT_T, 9> dm{};
    FLOAT_T ProductionHeightinCentimeter = 0;
    bool isSetProductionHeight = false;
    bool isSetCosine = false;
    bool isInit = true;
    int n_cosines;
    int n_energies;

private:
    // the reference evaluates sin/cos/sqrt through the double overloads and stores FLOAT_T
    static FLOAT_T sind(FLOAT_T x) { return FLOAT_T(std::sin(double(x))); }
    static FLOAT_T cosd(FLOAT_T x) { return FLOAT_T(std::cos(double(x))); }
    void set_u(int i, int j, FLOAT_T re, FLOAT_T im) {
        U(i, j).re = re;
        U(i, j).im = im;
    }
    void copy_state(const Propagator& o) {
        energyList = o.energyList;
        cosineList = o.cosineList;
        maxlayers = o.maxlayers;
        radii = o.radii;
        rhos = o.rhos;
        coslimit = o.coslimit;
        Mix_U = o.Mix_U;
        dm = o.dm;
        ProductionHeightinCentimeter = o.ProductionHeightinCentimeter;
        isSetCosine = o.isSetCosine;
        isSetProductionHeight = o.isSetProductionHeight;
        n_cosines = o.n_cosines;
        n_energies = o.n_energies;
    }
};

}  // namespace cudaprob3

#endif
