// clscalarfield.cpp -- CLScalarField over gfsx_p2g (see clscalarfield.h).
//
// Semantics kept from the reference (clscalarfield.cpp):
//  * initialize() returns false when no device is usable (:31-66);
//  * addPoints / addPointValues accumulate into the caller's arrays, as the
//    OpenCL path does (:1448-1449), with node (i,j,k) at offset + (i,j,k)*dx;
//  * the weighted overload fills the weight field too; the ScalarField
//    overloads use its weight field iff it is enabled (:269-301).
// The maximum-value threshold of addPoints (a chunk-culling heuristic of the
// OpenCL path, :992-1013) is not applied: points are always splatted.
// disableOpenCL() selects the reference's CPU path (:75-77, :499-509), which
// this drop-in does not carry: the flag is kept and reported truthfully, and
// a splat while it is cleared fails with std::runtime_error instead of
// silently running on the device.
#include "clscalarfield.h"

#include <cstdio>
#include <iostream>
#include <stdexcept>

#include "gfsx.h"

namespace {
void check(int err, const char *what) {
    if (err != GFSX_SUCCESS)
        throw std::runtime_error(std::string(what) + ": " + gfsx_get_error_message());
}
}  // namespace

CLScalarField::CLScalarField() {}

CLScalarField::~CLScalarField() {
    int err;
    if (_ctx) gfsx_destroy(_ctx, &err);
}

bool CLScalarField::initialize() {
    if (_ctx) return true;
    int err;
    _ctx = gfsx_create(0, &err);
    return err == GFSX_SUCCESS && _ctx != nullptr;
}

void CLScalarField::_splat(std::vector<vmath::vec3> &points, const float *values,
                           double radius, vmath::vec3 offset, double dx,
                           Array3d<float> *field, Array3d<float> *weight) {
    if (!_isOpenCLEnabled)
        throw std::runtime_error("CLScalarField: the accelerated path is disabled "
                                 "(disableOpenCL) and this drop-in has no CPU path; "
                                 "call enableOpenCL() or use ScalarField::addPointValue");
    if (!_ctx && !initialize())
        throw std::runtime_error("CLScalarField: no HIP device (" +
                                 std::string(gfsx_get_error_message()) + ")");
    if (weight && (weight->width != field->width || weight->height != field->height ||
                   weight->depth != field->depth))
        throw std::invalid_argument("CLScalarField: scalar and weight fields differ in size");
    const float off[3] = {offset.x, offset.y, offset.z};
    int err;
    // vmath::vec3 is three packed floats, so the vector is an (n, 3) array
    gfsx_p2g(_ctx, reinterpret_cast<const float *>(points.data()), values,
             (int64_t)points.size(), radius, off, dx, field->width, field->height,
             field->depth, field->getRawArray(), weight ? weight->getRawArray() : nullptr,
             &err);
    check(err, "gfsx_p2g");
}

void CLScalarField::addPoints(std::vector<vmath::vec3> &points, double radius,
                              vmath::vec3 offset, double dx, Array3d<float> *field) {
    _splat(points, nullptr, radius, offset, dx, field, nullptr);
}

void CLScalarField::addPoints(std::vector<vmath::vec3> &points, double radius,
                              vmath::vec3 offset, double dx, ScalarField &isfield) {
    addPoints(points, radius, offset, dx, isfield.getPointerToScalarField());
}

void CLScalarField::addPoints(std::vector<vmath::vec3> &points, ScalarField &isfield) {
    addPoints(points, isfield.getPointRadius(), isfield.getOffset(), isfield.getCellSize(),
              isfield.getPointerToScalarField());
}

void CLScalarField::addPointValues(std::vector<vmath::vec3> &points,
                                   std::vector<float> &values, double radius,
                                   vmath::vec3 offset, double dx, Array3d<float> *field) {
    if (points.size() != values.size())
        throw std::invalid_argument("CLScalarField: points and values differ in length");
    _splat(points, values.data(), radius, offset, dx, field, nullptr);
}

void CLScalarField::addPointValues(std::vector<vmath::vec3> &points,
                                   std::vector<float> &values, double radius,
                                   vmath::vec3 offset, double dx,
                                   Array3d<float> *scalarfield,
                                   Array3d<float> *weightfield) {
    if (points.size() != values.size())
        throw std::invalid_argument("CLScalarField: points and values differ in length");
    _splat(points, values.data(), radius, offset, dx, scalarfield, weightfield);
}

void CLScalarField::addPointValues(std::vector<vmath::vec3> &points,
                                   std::vector<float> &values, double radius,
                                   vmath::vec3 offset, double dx, ScalarField &isfield) {
    Array3d<float> *w = isfield.isWeightFieldEnabled() ? isfield.getPointerToWeightField()
                                                        : nullptr;
    addPointValues(points, values, radius, offset, dx, isfield.getPointerToScalarField(), w);
}

void CLScalarField::addPointValues(std::vector<vmath::vec3> &points,
                                   std::vector<float> &values, ScalarField &isfield) {
    addPointValues(points, values, isfield.getPointRadius(), isfield.getOffset(),
                   isfield.getCellSize(), isfield);
}

void CLScalarField::setMaxScalarFieldValueThreshold(float val) {
    _isMaxScalarFieldValueThresholdSet = true;
    _maxScalarFieldValueThreshold = val;
}
void CLScalarField::setMaxScalarFieldValueThreshold() { _isMaxScalarFieldValueThresholdSet = false; }
bool CLScalarField::isMaxScalarFieldValueThresholdSet() { return _isMaxScalarFieldValueThresholdSet; }
double CLScalarField::getMaxScalarFieldValueThreshold() { return _maxScalarFieldValueThreshold; }

// The HIP device is the only device; the preference setters have nothing to
// choose between.
void CLScalarField::setDevicePreference(std::string) {}
void CLScalarField::setDevicePreferenceGPU() {}
void CLScalarField::setDevicePreferenceCPU() {}

std::string CLScalarField::getDeviceInfo() {
    if (!_ctx) return "";
    int err;
    return gfsx_get_device_info(_ctx, &err);
}
void CLScalarField::printDeviceInfo() { std::cout << getDeviceInfo() << std::endl; }
// The splat every addPoints / addPointValues call runs (gfsx_p2g): a
// counting sort of the points by node cell, then one gather per node over
// the cells within the radius.
std::string CLScalarField::getKernelInfo() {
    return _ctx ? "gfsx_p2g: cell-sorted gather P2G (k_p2g_gather, gfx950)" : "";
}
void CLScalarField::printKernelInfo() { std::cout << getKernelInfo() << std::endl; }
bool CLScalarField::isUsingGPU() { return _ctx != nullptr; }
bool CLScalarField::isUsingCPU() { return false; }
void CLScalarField::disableOpenCL() { _isOpenCLEnabled = false; }
void CLScalarField::enableOpenCL() { _isOpenCLEnabled = true; }
bool CLScalarField::isOpenCLEnabled() { return _isOpenCLEnabled; }
int CLScalarField::getKernelWorkLoadSize() { return _kernelWorkLoadSize; }
void CLScalarField::setKernelWorkLoadSize(int n) { _kernelWorkLoadSize = n; }
