// particleadvector.h -- drop-in replacement for GridFluidSim3D's
// ParticleAdvector (particleadvector.h:65-110 of the reference): the same
// class name and public interface, implemented over the gfsx C ABI
// (include/gfsx.h) on a HIP device instead of OpenCL.  Needs vmath.h and
// macvelocityfield.h from the reference tree and links libgfsx.so.
#ifndef GFSX_DROPIN_PARTICLEADVECTOR_H
#define GFSX_DROPIN_PARTICLEADVECTOR_H

#include <string>
#include <vector>

#include "macvelocityfield.h"
#include "vmath.h"

struct gfsx_ctx;

class ParticleAdvector {
public:
    ParticleAdvector();
    ~ParticleAdvector();
    ParticleAdvector(const ParticleAdvector &) = delete;
    ParticleAdvector &operator=(const ParticleAdvector &) = delete;

    bool initialize();
    void setDevicePreference(std::string devtype);
    void setDevicePreferenceGPU();
    void setDevicePreferenceCPU();

    void printDeviceInfo();
    std::string getDeviceInfo();
    void printKernelInfo();
    std::string getKernelInfo();
    bool isUsingGPU();
    bool isUsingCPU();
    void disableOpenCL();
    void enableOpenCL();
    bool isOpenCLEnabled();
    int getKernelWorkLoadSize();
    void setKernelWorkLoadSize(int n);

    void advectParticlesRK4(std::vector<vmath::vec3> &particles,
                            MACVelocityField *vfield, double dt,
                            std::vector<vmath::vec3> &output);
    void advectParticlesRK3(std::vector<vmath::vec3> &particles,
                            MACVelocityField *vfield, double dt,
                            std::vector<vmath::vec3> &output);
    void advectParticlesRK2(std::vector<vmath::vec3> &particles,
                            MACVelocityField *vfield, double dt,
                            std::vector<vmath::vec3> &output);
    void advectParticlesRK1(std::vector<vmath::vec3> &particles,
                            MACVelocityField *vfield, double dt,
                            std::vector<vmath::vec3> &output);

    void tricubicInterpolate(std::vector<vmath::vec3> &particles,
                             MACVelocityField *vfield,
                             std::vector<vmath::vec3> &output);
    // overwrites particles with the output
    void tricubicInterpolate(std::vector<vmath::vec3> &particles,
                             MACVelocityField *vfield);

    // The G2P stencil used by every method above: GFSX_TRICUBIC (the
    // reference's, the default) or GFSX_TRILINEAR (the stencil of
    // MACVelocityField::evaluateVelocityAtPositionLinear).  No counterpart
    // in the reference; anything else throws std::invalid_argument.
    int stencil() const { return _stencil; }
    void setStencil(int stencil);

private:
    void _advect(int order, std::vector<vmath::vec3> &particles,
                 MACVelocityField *vfield, double dt,
                 std::vector<vmath::vec3> &output);
    void _ensure();

    gfsx_ctx *_ctx = nullptr;
    int _stencil = 1;               // GFSX_TRICUBIC
    bool _isOpenCLEnabled = true;   // the accelerated (here: HIP) path, as in the reference
    int _kernelWorkLoadSize = 1000;
};

#endif
