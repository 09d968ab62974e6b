// clscalarfield.h -- drop-in replacement for GridFluidSim3D's CLScalarField
// (clscalarfield.h:67-125 of the reference): the same class name and public
// interface, implemented over the gfsx C ABI (include/gfsx.h) on a HIP
// device instead of OpenCL.  Compile it in the reference tree in place of
// the original clscalarfield.{h,cpp}; it needs vmath.h, array3d.h and
// scalarfield.h from that tree and links libgfsx.so.
#ifndef GFSX_DROPIN_CLSCALARFIELD_H
#define GFSX_DROPIN_CLSCALARFIELD_H

#include <string>
#include <vector>

#include "array3d.h"
#include "scalarfield.h"
#include "vmath.h"

struct gfsx_ctx;

class CLScalarField {
public:
    CLScalarField();
    ~CLScalarField();
    CLScalarField(const CLScalarField &) = delete;
    CLScalarField &operator=(const CLScalarField &) = delete;

    bool initialize();
    void addPoints(std::vector<vmath::vec3> &points, double radius,
                   vmath::vec3 offset, double dx, Array3d<float> *field);
    void addPoints(std::vector<vmath::vec3> &points, double radius,
                   vmath::vec3 offset, double dx, ScalarField &field);
    void addPoints(std::vector<vmath::vec3> &points, ScalarField &field);

    void addPointValues(std::vector<vmath::vec3> &points,
                        std::vector<float> &values, double radius,
                        vmath::vec3 offset, double dx, Array3d<float> *field);
    void addPointValues(std::vector<vmath::vec3> &points,
                        std::vector<float> &values, double radius,
                        vmath::vec3 offset, double dx,
                        Array3d<float> *scalarfield,
                        Array3d<float> *weightfield);
    void addPointValues(std::vector<vmath::vec3> &points,
                        std::vector<float> &values, double radius,
                        vmath::vec3 offset, double dx, ScalarField &field);
    void addPointValues(std::vector<vmath::vec3> &points,
                        std::vector<float> &values, ScalarField &field);

    void setMaxScalarFieldValueThreshold(float val);
    void setMaxScalarFieldValueThreshold();
    bool isMaxScalarFieldValueThresholdSet();
    double getMaxScalarFieldValueThreshold();

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

private:
    void _splat(std::vector<vmath::vec3> &points, const float *values,
                double radius, vmath::vec3 offset, double dx,
                Array3d<float> *field, Array3d<float> *weight);

    gfsx_ctx *_ctx = nullptr;
    bool _isOpenCLEnabled = true;   // the accelerated (here: HIP) path, as in the reference
    bool _isMaxScalarFieldValueThresholdSet = false;
    float _maxScalarFieldValueThreshold = 1.0f;
    int _kernelWorkLoadSize = 1000;
};

#endif
