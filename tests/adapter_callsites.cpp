// Compile check of the drop-in accelerator: HiDEMLib's own call expressions
// on its accelerator object (HiDEMLib.h:265, 332, 347, 603-607, 620-622,
// 814-818) and the sample main's functor constructions (MainDem.cpp:297-314),
// written as HiDEMLib writes them, against CombinedAccelHip.  The data
// handlers, execution policy, communicator and ManVec packs are stand-ins
// with the reference's names; the accelerator ignores them.
//
// Built with -DHDM_WRONG_FUNCTOR, a lambda in place of the default link
// functor must fail to compile.
#include <cstdio>
#include <optional>
#include <string>
#include <tuple>

#include "hidem/CombinedAccelHip.h"

using namespace hidem;

struct ParticleData {};
struct LinkData {};
struct OffloadExec {};
struct PForceComm {};
struct Distribution {
    PForceComm pForceComm;
};
template <class... T> std::tuple<T *...> packVecPtrs(T &...v) { return std::tuple<T *...>(&v...); }
template <class... A> auto make4Pack(A... a) { return std::make_tuple(a...); }
template <class... A> auto make3Pack(A... a) { return std::make_tuple(a...); }

struct HiDEMLibShape {
    CombinedAccelHip *accel = nullptr;
    ParticleData *pDataPtr = nullptr;
    LinkData *lDataPtr = nullptr;
    Distribution *distr = nullptr;
    OffloadExec execType;
    double searchDistance = 0.85, interactDist = 0.14;
    bool writeHelperStructures = false, linkRegenEnabled = false;
    std::optional<std::string> outWorkingPath;
    int currentTimestep = 0;
    int getLocalRank() const { return 0; }

    template <class NFunc, class LFunc, class... N, class... L, class... S, class... F>
    void computePartForce(NFunc neighbourForceFunc, LFunc linkForceFunc, const std::tuple<N...> neighArgs,
                          const std::tuple<L...> linkArgs, std::tuple<S...> linkSideArgs, std::tuple<F...> forceArgs) {
        auto fullArgs = make4Pack(neighArgs, linkArgs, linkSideArgs, forceArgs);
        accel->iterateCombinedPre(execType, *pDataPtr, *lDataPtr, neighbourForceFunc, linkForceFunc, fullArgs);
        accel->iterateCombinedPost(execType, *pDataPtr, *lDataPtr, neighbourForceFunc, linkForceFunc, fullArgs);
    }

    template <class LinkFun, class... N, class... S, class... L>
    void linkUpdate(LinkFun breakLinks, std::tuple<N...> neighArgs, std::tuple<S...> linkSideArgs,
                    std::tuple<L...> linkArgs) {
        auto fullArgs = make3Pack(neighArgs, linkSideArgs, linkArgs);
        accel->iterateLinks(execType, *pDataPtr, *lDataPtr, breakLinks, fullArgs);
    }

    void gridRegenerate() {
        accel->generatePrePostCommIndexes(*pDataPtr, *lDataPtr, distr->pForceComm);
        accel->generateNeighbours(*pDataPtr, *lDataPtr, searchDistance, interactDist);
        if (writeHelperStructures)
            accel->grid.toVTKHDF(outWorkingPath.value_or(".") + "/grid_" + std::to_string(getLocalRank())
                + "_" + std::to_string(currentTimestep) + ".vtkhdf");
        accel->generateLinks(*pDataPtr, *lDataPtr, searchDistance, interactDist);
        accel->lAccel.sortSideData(*pDataPtr, *lDataPtr);
        accel->generateIndexes(*pDataPtr, *lDataPtr, searchDistance, interactDist);
    }

    void setupAccel() {
        accel->lAccel.keepInactive = linkRegenEnabled;
        accel->generatePrePostCommIndexes(*pDataPtr, *lDataPtr, distr->pForceComm);
        accel->generateNeighbours(*pDataPtr, *lDataPtr, searchDistance, interactDist);
        accel->generateLinks(*pDataPtr, *lDataPtr, searchDistance, interactDist);
        accel->generateIndexes(*pDataPtr, *lDataPtr, searchDistance, interactDist);
    }
};

// The sample loop's calls (MainDem.cpp:297-314) with its ManVec stand-ins.
void sample_step(HiDEMLibShape &hiDem) {
    double invDt = 1e4, attractDistance = 0.0, attractScaleForce = 0.0, damping = 0.0, dampingRotation = 0.0,
           scaleE = 1e9, E = 1e9, W2 = 1.0, W4 = 1.0, G = 5e8, bendBreak = 1.0, maxLoad = 3e-4;
    double xDiff = 0, yDiff = 0, zDiff = 0, uDiff = 0, vDiff = 0, wDiff = 0, radius = 0;
    double CT0 = 0, CT1 = 0, CT2 = 0, CT3 = 0, CT4 = 0, CT5 = 0;
    double xForce = 0, yForce = 0, zForce = 0, uForce = 0, vForce = 0, wForce = 0;
    double xCord = 0, yCord = 0, zCord = 0, xTheta = 0, yTheta = 0, zTheta = 0, initialLen = 0;
    unsigned char state = 1;
    hiDem.gridRegenerate();
#ifndef HDM_WRONG_FUNCTOR
    hiDem.computePartForce(HiDEMDefNeighForceFunc(invDt, attractDistance, attractScaleForce, damping, dampingRotation, scaleE),
        HiDEMDefLinkForceFunc(invDt, damping, dampingRotation, E, W2, W4, G),
#else
    hiDem.computePartForce(HiDEMDefNeighForceFunc(invDt, attractDistance, attractScaleForce, damping, dampingRotation, scaleE),
        [=](double &mask) { mask = G; },
#endif
                           packVecPtrs(xDiff, yDiff, zDiff, uDiff, vDiff, wDiff, radius),
                           packVecPtrs(),
                           packVecPtrs(CT0, CT1, CT2, CT3, CT4, CT5),
                           packVecPtrs(xForce, yForce, zForce, uForce, vForce, wForce));
    hiDem.linkUpdate(HiDEMDefLinkBreakFunc(bendBreak, maxLoad),
                     packVecPtrs(xCord, yCord, zCord, xTheta, yTheta, zTheta),
                     packVecPtrs(),
                     packVecPtrs(state, initialLen));
}

// Run on a GPU: two particles 1 m apart joined by one link, the second
// moving 1 um along the link in the step; prints the axial forces, which the
// beam law makes equal and opposite (E W^2 / L * 1e-6 = 1000 N).
int main() {
    ParticleData pData;
    LinkData lData;
    Distribution distr;
    CombinedAccelHip accel(HDM_PREC_FP64, 0);
    double x[2] = {10.0, 11.0}, y[2] = {10.0, 10.0}, z[2] = {10.0, 10.0}, zero[2] = {0.0, 0.0};
    double xd[2] = {0.0, 1e-6}, r[2] = {0.5, 0.5};
    hdm_particles_host p{};
    p.n = 2;
    p.xCord = x; p.yCord = y; p.zCord = z;
    p.xTheta = zero; p.yTheta = zero; p.zTheta = zero;
    p.xDiff = xd; p.yDiff = zero; p.zDiff = zero; p.uDiff = zero; p.vDiff = zero; p.wDiff = zero;
    p.radius = r;
    int32_t a[1] = {0}, b[1] = {1};
    uint8_t st[1] = {1};
    double len[1] = {1.0}, ct[6][2] = {};
    hdm_links_host l{};
    l.n = 1;
    l.localPartIndA = a; l.localPartIndB = b; l.state = st; l.initialLen = len;
    for (int k = 0; k < 6; k++) l.CT[k] = ct[k];
    accel.uploadParticles(p);
    accel.uploadLinks(l);
    HiDEMLibShape h;
    h.accel = &accel;
    h.pDataPtr = &pData;
    h.lDataPtr = &lData;
    h.distr = &distr;
    h.setupAccel();
    sample_step(h);
    double fx[2] = {0.0, 0.0};
    hdm_particles_host q{};
    q.n = 2;
    q.xForce = fx;
    accel.downloadParticles(q);
    std::printf("xForce %.17g %.17g\n", fx[0], fx[1]);
    return 0;
}
