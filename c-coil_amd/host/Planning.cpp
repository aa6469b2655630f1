// Quadrature planning: PrecisionFactor, PrecisionArguments, CoilPairArguments.
// Integer outputs must equal the reference's bit for bit (checked against the oracle in tests); the
// algorithm is the reference's:
//   PrecisionFactor            src/Coil/PrecisionArguments/PrecisionFactor.cxx:13-21
//   PrecisionArguments ctor    src/Coil/PrecisionArguments/PrecisionArguments.cxx:29-54
//   coil planning              PrecisionArguments.cxx:69-185 (CPU and GPU), :188-288 (GPU secondary)
//   pair planning              src/Coil/PrecisionArguments/CoilPairArguments.cxx:20-255
#include "Coil.h"

#include <cmath>
#include <cstdlib>
#include <sstream>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxOrder = 100;          // Legendre::maxLegendreOrder
constexpr int kDefaultOrder = 20;       // PrecisionArguments.cxx:13
constexpr int kBaseIncrements = 10;     // g_baseLayerIncrementsCPU / GPU (PrecisionGlobalVars.h:13-14)
constexpr double kPrimWeight = 1.414213562;  // g_primAngularWeightModifier / g_primLinearWeightModifier
constexpr double kSecWeight = 1.0;           // g_secAngularWeightModifier / g_secLinearWeightModifier

int g_gpu_increments = -1;

int layersToBlocks(int layer) { return (layer - 1) / kMaxOrder + 1; }
int layersToIncrements(int layer)
{
    const int blocks = layersToBlocks(layer);
    return blocks > 1 ? layer / blocks + 1 : layer;
}

// Balances and spills the total increment budget over the given (layer, measure) components: the do/while
// loop shared by both planners. capAllowed(layer) says whether a layer may be capped at the GPU limit.
template <typename CapAllowed>
void balanceWithSpill(int total, std::vector<std::pair<int, double>> components, bool gpu, int cap,
                      CapAllowed capAllowed, int *layer)
{
    std::vector<std::pair<int, int>> counts;
    bool spilled;
    do {
        counts = CoilPairArguments::balanceIncrements(total, components);
        spilled = false;
        if (gpu) {
            for (size_t i = 0; i < counts.size(); ++i) {
                if (counts[i].second > cap && capAllowed(counts[i].first)) {
                    layer[counts[i].first] = cap;
                    components.erase(components.begin() + (long)i);
                    counts.erase(counts.begin() + (long)i);
                    total = (int)std::ceil((double)total / (double)cap);
                    spilled = true;
                    break;
                }
            }
        }
    } while (spilled && !components.empty());
    for (const auto &c : counts) layer[c.first] = c.second;
}

}  // namespace

namespace ccoil {

int gpuIncrements()
{
    if (g_gpu_increments < 0) {
        int v = 80;
        if (const char *e = std::getenv("CCOIL_GPU_INCREMENTS")) {
            int x = std::atoi(e);
            if (x >= 1 && x <= kMaxOrder) v = x;
        }
        g_gpu_increments = v;
    }
    return g_gpu_increments;
}

void setGpuIncrements(int increments)
{
    if (increments < 1 || increments > kMaxOrder) increments = 80;
    g_gpu_increments = increments;
}

}  // namespace ccoil

PrecisionFactor::PrecisionFactor() : PrecisionFactor(5.0) {}

PrecisionFactor::PrecisionFactor(double rp)
{
    relativePrecision = (rp < 1.0 || rp > 15.0) ? 5.0 : rp;
}

PrecisionFactor::operator std::string() const
{
    std::ostringstream o;
    o << "PrecisionFactor(relative_precision=" << relativePrecision << ")";
    return o.str();
}

PrecisionArguments::PrecisionArguments()
    : PrecisionArguments(1, 1, 1, kDefaultOrder, kDefaultOrder, kDefaultOrder) {}

PrecisionArguments::PrecisionArguments(int ab, int tb, int lb, int ai, int ti, int li)
    : angularBlocks(ab), thicknessBlocks(tb), lengthBlocks(lb), angularIncrements(ai), thicknessIncrements(ti),
      lengthIncrements(li)
{
    // Silent fallbacks, as the reference does (out-of-range increments -> 20, blocks < 1 -> 1); the length
    // block check tests lengthIncrements in the reference (PrecisionArguments.cxx:52-53) and does so here.
    if (ai > kMaxOrder || ai < 1) angularIncrements = kDefaultOrder;
    if (ti > kMaxOrder || ti < 1) thicknessIncrements = kDefaultOrder;
    if (li > kMaxOrder || li < 1) lengthIncrements = kDefaultOrder;
    if (ab < 1) angularBlocks = 1;
    if (tb < 1) thicknessBlocks = 1;
    if (li < 1) lengthBlocks = 1;
}

PrecisionArguments PrecisionArguments::getCoilPrecisionArgumentsCPU(const Coil &coil, PrecisionFactor pf)
{
    return calculatePrecisionArguments(coil, pf, false);
}

PrecisionArguments PrecisionArguments::getCoilPrecisionArgumentsGPU(const Coil &coil, PrecisionFactor pf)
{
    return calculatePrecisionArguments(coil, pf, true);
}

PrecisionArguments PrecisionArguments::calculatePrecisionArguments(const Coil &coil, PrecisionFactor pf, bool gpu)
{
    const int cap = ccoil::gpuIncrements();
    int layer[3] = {1, 1, 1};
    const double m[3] = {kPrimWeight * std::sqrt(kPi * (coil.getInnerRadius() + 0.5 * coil.getThickness())),
                         kPrimWeight * std::sqrt(coil.getThickness()), kPrimWeight * std::sqrt(coil.getLength())};
    std::vector<std::pair<int, double>> comps{{0, m[0]}};
    int total = kBaseIncrements;
    const CoilType t = coil.getCoilType();
    if (t == CoilType::RECTANGULAR || t == CoilType::FLAT) {
        total *= kBaseIncrements;
        comps.emplace_back(1, m[1]);
    }
    total = (int)((double)total * std::pow(2, pf.relativePrecision - 1.0));
    balanceWithSpill(total, comps, gpu, cap, [](int) { return true; }, layer);
    if (layer[2] == 1) layer[2] = (int)std::round((double)layer[0] * m[2] / m[0]);
    if (layer[2] == 0) layer[2] = 1;
    if (gpu && layer[2] > cap) layer[2] = cap;
    return PrecisionArguments(layersToBlocks(layer[0]), layersToBlocks(layer[1]), layersToBlocks(layer[2]),
                              layersToIncrements(layer[0]), layersToIncrements(layer[1]), layersToIncrements(layer[2]));
}

PrecisionArguments PrecisionArguments::getSecondaryCoilPrecisionArgumentsGPU(const Coil &coil, PrecisionFactor pf)
{
    const int cap = ccoil::gpuIncrements();
    int layer[3] = {1, 1, 1};
    const double m[3] = {kPrimWeight * std::sqrt(kPi * (coil.getInnerRadius() + 0.5 * coil.getThickness())),
                         kPrimWeight * std::sqrt(coil.getThickness()), kPrimWeight * std::sqrt(coil.getLength())};
    std::vector<std::pair<int, double>> comps{{0, m[0]}};
    int total = kBaseIncrements;
    switch (coil.getCoilType()) {
        case CoilType::RECTANGULAR:
            total *= kBaseIncrements * kBaseIncrements;
            comps.emplace_back(1, m[1]);
            comps.emplace_back(2, m[2]);
            break;
        case CoilType::THIN:
            total *= kBaseIncrements;
            comps.emplace_back(2, m[2]);
            break;
        case CoilType::FLAT:
            total *= kBaseIncrements;
            comps.emplace_back(1, m[1]);
            break;
        case CoilType::FILAMENT:
            break;
    }
    total = (int)((double)total * std::pow(2, pf.relativePrecision - 1.0));
    balanceWithSpill(total, comps, true, cap, [](int) { return true; }, layer);
    if (layer[2] == 1) layer[2] = (int)std::round((double)layer[0] * m[2] / m[0]);
    if (layer[2] == 0) layer[2] = 1;
    if (layer[2] > cap) layer[2] = cap;
    return PrecisionArguments(1, 1, 1, layer[0], layer[1], layer[2]);
}

PrecisionArguments::operator std::string() const
{
    std::ostringstream o;
    o << "PrecisionArguments(angular_block_count=" << angularBlocks << ", thickness_block_count=" << thicknessBlocks
      << ", length_block_count=" << lengthBlocks << ", angular_increment_count=" << angularIncrements
      << ", thickness_increment_count=" << thicknessIncrements << ", length_increment_count=" << lengthIncrements
      << ")";
    return o.str();
}

CoilPairArguments::CoilPairArguments() : CoilPairArguments(PrecisionArguments(), PrecisionArguments()) {}

CoilPairArguments::CoilPairArguments(const PrecisionArguments &p, const PrecisionArguments &s)
    : primaryPrecision(p), secondaryPrecision(s) {}

std::vector<std::pair<int, int>> CoilPairArguments::balanceIncrements(int total,
                                                                      const std::vector<std::pair<int, double>> &comps)
{
    std::vector<std::pair<int, int>> out;
    out.reserve(comps.size());
    if (comps.size() < 2) {
        out.emplace_back(comps[0].first, total);
        return out;
    }
    double measure = 1.0;
    for (const auto &c : comps) measure *= c.second;
    measure /= (double)total;
    const double linear = std::pow(measure, 1.0 / (double)comps.size());
    for (const auto &c : comps) out.emplace_back(c.first, (int)std::round(c.second / linear));
    return out;
}

CoilPairArguments CoilPairArguments::getAppropriateCoilPairArguments(const Coil &prim, const Coil &sec,
                                                                     PrecisionFactor pf, ComputeMethod method,
                                                                     bool zAxisCase, bool pureGPU)
{
    const int cap = ccoil::gpuIncrements();
    const bool gpu = method == GPU;
    int layer[6] = {1, 1, 1, 1, 1, 1};
    const double m[6] = {kPrimWeight * std::sqrt(kPi * (prim.getInnerRadius() + 0.5 * prim.getThickness())),
                         kPrimWeight * std::sqrt(prim.getThickness()),
                         kPrimWeight * std::sqrt(prim.getLength()),
                         kSecWeight * std::sqrt(2 * kPi * (sec.getInnerRadius() + 0.5 * sec.getThickness())),
                         kSecWeight * std::sqrt(sec.getThickness()),
                         kSecWeight * std::sqrt(sec.getLength())};
    std::vector<std::pair<int, double>> comps{{0, m[0]}};
    int total = kBaseIncrements;
    const CoilType pt = prim.getCoilType(), st = sec.getCoilType();
    if (pt == CoilType::RECTANGULAR || pt == CoilType::FLAT) {
        total *= kBaseIncrements;
        comps.emplace_back(1, m[1]);
    }
    if (zAxisCase) {
        if (st == CoilType::RECTANGULAR || st == CoilType::FLAT) {
            total *= kBaseIncrements;
            comps.emplace_back(4, m[4]);
        }
    } else {
        comps.emplace_back(3, m[3]);
        total *= kBaseIncrements;
        switch (st) {
            case CoilType::RECTANGULAR:
                total *= kBaseIncrements * kBaseIncrements;
                comps.emplace_back(4, m[4]);
                comps.emplace_back(5, m[5]);
                break;
            case CoilType::THIN:
                total *= kBaseIncrements;
                comps.emplace_back(5, m[5]);
                break;
            case CoilType::FLAT:
                total *= kBaseIncrements;
                comps.emplace_back(4, m[4]);
                break;
            case CoilType::FILAMENT:
                break;
        }
    }
    total = (int)((double)total * std::pow(2, pf.relativePrecision - 1.0));
    balanceWithSpill(total, comps, gpu, cap, [&](int l) { return l < 3 || pureGPU; }, layer);
    for (int i = 1; i < 6; ++i) {
        if (layer[i] == 1) layer[i] = (int)std::round((double)layer[0] * m[i] / m[0]);
        if (layer[i] == 0) layer[i] = 1;
        if (gpu && layer[i] >= cap) layer[i] = cap;
    }
    PrecisionArguments p(layersToBlocks(layer[0]), layersToBlocks(layer[1]), layersToBlocks(layer[2]),
                         layersToIncrements(layer[0]), layersToIncrements(layer[1]), layersToIncrements(layer[2]));
    PrecisionArguments s(layersToBlocks(layer[3]), layersToBlocks(layer[4]), layersToBlocks(layer[5]),
                         layersToIncrements(layer[3]), layersToIncrements(layer[4]), layersToIncrements(layer[5]));
    return CoilPairArguments(p, s);
}

CoilPairArguments::operator std::string() const
{
    std::ostringstream o;
    o << "CoilPairArguments(primary_precision=" << std::string(primaryPrecision)
      << ", secondary_precision=" << std::string(secondaryPrecision) << ")";
    return o.str();
}
