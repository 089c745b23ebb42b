// ref_leaf.cpp — exposes the reference's own leaf PRNG/hash/frame/Box-Muller code
// (compiled from the reference sources where they lie) through a C ABI, so the
// oracle's restatement can be checked against it.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/build_ref.sh and
// loaded only by tests/ (and the fixture script under tests/golden/).
//
// Reference code used (not copied; #included from the reference tree):
//   src/core/math/MathUtil.hpp:179-228       xxhash32, Bernoulli
//   src/core/sampling/UniformSampler.hpp     PCG32 sampler, next1D/next3D/set_state
//   src/core/math/BitManip.hpp:47-50         normalizedUint
//   src/core/math/TangentFrame.hpp:74-85     TangentFrame(n)
//   src/core/sampling/Gaussian.cpp:22-34,89-104  rand_normal_2, sample_standard_normal
//   src/core/math/SdfFunctionsQueries.cpp        SdfFunctions::evaluateJet, eval (float API)
//   src/core/math/Mat4f.{hpp,cpp}                rotYXZ, invert, transformVector
//   src/core/math/SdfFunctions.cpp:84-183        random3, simplex3d, fbm (the ProceduralNoise fields)
#include "math/Mat4f.hpp"
#include "math/MathUtil.hpp"
#include "math/SdfFunctions.hpp"
#include "math/TangentFrame.hpp"
#include "sampling/Gaussian.hpp"
#include "sampling/UniformSampler.hpp"

#include <cstdint>

using namespace Tungsten;

namespace Tungsten {
Vec3f random3(Vec3f c);
float simplex3d(Vec3f p);
}

extern "C" {

// random3, simplex3d and fbm of SdfFunctions.cpp at n points (xyz triples).
void ref_random3(const float *p, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) {
        Vec3f r = random3(Vec3f(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
        out[3 * i] = r.x(); out[3 * i + 1] = r.y(); out[3 * i + 2] = r.z();
    }
}
void ref_simplex3d(const float *p, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = simplex3d(Vec3f(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
}
void ref_fbm(const double *p, int64_t n, int octaves, double *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = fbm(Vec3d(p[3 * i], p[3 * i + 1], p[3 * i + 2]), octaves);
}

uint32_t ref_xxhash32_1(uint32_t p) { return MathUtil::xxhash32(uint(p)); }
uint32_t ref_xxhash32_2(uint32_t x, uint32_t y) { return MathUtil::xxhash32(Vec2u(x, y)); }
uint32_t ref_xxhash32_4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    return MathUtil::xxhash32(Vec4u(x, y, z, w));
}

// The statements of SparseConvolutionNoiseRealization::cell3D
// (SparseConvolutionNoise.cpp:433-442) that touch the sampler, written with the
// reference's own UniformSampler and Vec3f(next3D()).
void ref_cell3d_impulses(uint32_t i, uint32_t j, uint32_t k, uint32_t seed, uint32_t n, float *out) {
    UniformSampler sampler;
    sampler.set_state(MathUtil::xxhash32(Vec4u(k, j, i, seed)) + 1u);
    for (uint32_t q = 0; q < n; ++q) {
        Vec3f p_i = Vec3f(sampler.next3D());
        float w_i = MathUtil::Bernoulli(sampler.next1D(), -1.f, 1.f, 0.5f);
        out[4 * q + 0] = p_i.x();
        out[4 * q + 1] = p_i.y();
        out[4 * q + 2] = p_i.z();
        out[4 * q + 3] = w_i;
    }
}

// cell1D (SparseConvolutionNoise.cpp:466-476) sampler statements.
void ref_cell1d_impulses(uint32_t i, uint32_t seed, uint32_t n, float *out) {
    UniformSampler sampler;
    sampler.set_state(MathUtil::xxhash32(Vec2u(i, seed)) + 1u);
    for (uint32_t q = 0; q < n; ++q) {
        float t_i = sampler.next1D();
        float w_i = MathUtil::Bernoulli(sampler.next1D(), -1.f, 1.f, 0.5f);
        out[2 * q] = t_i;
        out[2 * q + 1] = w_i;
    }
}

void ref_standard_normals(uint64_t state, int npairs, double *out) {
    UniformSampler sampler;
    sampler.set_state(state);
    for (int m = 0; m < npairs; ++m) {
        Eigen::VectorXd z = sample_standard_normal(2, sampler);
        out[2 * m] = z[0];
        out[2 * m + 1] = z[1];
    }
}

void ref_tangent_frame(const float n[3], float out9[9]) {
    TangentFrame f(Vec3f(n[0], n[1], n[2]));
    out9[0] = f.tangent.x(); out9[1] = f.tangent.y(); out9[2] = f.tangent.z();
    out9[3] = f.bitangent.x(); out9[4] = f.bitangent.y(); out9[5] = f.bitangent.z();
    out9[6] = f.normal.x(); out9[7] = f.normal.y(); out9[8] = f.normal.z();
}

// Raw PCG output words after set_state(state): the sampler's full sequence.
void ref_pcg_words(uint64_t state, int n, uint32_t *out) {
    UniformSampler sampler;
    sampler.set_state(state);
    for (int q = 0; q < n; ++q) out[q] = sampler.nextI();
}

// SdfFunctions::evaluateJet at a local point: value, gradient into g.
double ref_sdf_jet(int fn, const double p[3], double g[3]) {
    const SdfJet q = SdfFunctions::evaluateJet(SdfFunctions::Function(fn), Eigen::Vector3d(p[0], p[1], p[2]));
    g[0] = q.gradient.x(); g[1] = q.gradient.y(); g[2] = q.gradient.z();
    return q.value;
}

// SdfFunctions::eval, the float API.
float ref_sdf_eval(int fn, const float p[3]) {
    int mat = 0;
    return SdfFunctions::eval(SdfFunctions::Function(fn), Vec3f(p[0], p[1], p[2]), mat);
}

// The statements of ProceduralMean::prepareSdfQueries and evaluateJet
// (ProceduralMeanSdf.cpp:11-16, 32-38) on the reference's Eigen, which fix the order of
// the 3x3 double products; inv16 is _invConfigTransform, row-major.
double ref_proc_mean_jet(int fn, const float inv16[16], float minv, float scale, float offset, const double a[3],
                         double g[3]) {
    Eigen::Matrix3d w2l;
    Eigen::Vector3d tr;
    for (int i = 0; i < 3; ++i) {
        tr[i] = inv16[4 * i + 3];
        for (int j = 0; j < 3; ++j) w2l(i, j) = inv16[4 * i + j];
    }
    const Eigen::Vector3d p(a[0], a[1], a[2]);
    const SdfJet q = SdfFunctions::evaluateJet(SdfFunctions::Function(fn), w2l * p + tr);
    const double value = q.value * double(scale) + double(offset);
    Eigen::Vector3d grad = Eigen::Vector3d::Zero();
    double out = value;
    if (value < minv || scale == 0) out = std::max(value, double(minv));
    else grad = double(scale) * w2l.transpose() * q.gradient;
    g[0] = grad.x(); g[1] = grad.y(); g[2] = grad.z();
    return out;
}

// Mat4f::rotYXZ(rot), Mat4f::invert and Mat4f * Vec3f, row-major 4x4.
void ref_mat4f_rot_yxz(const float rot[3], float out16[16]) {
    const Mat4f m = Mat4f::rotYXZ(Vec3f(rot[0], rot[1], rot[2]));
    for (int i = 0; i < 16; ++i) out16[i] = m[i];
}
void ref_mat4f_invert(const float m16[16], float out16[16]) {
    Mat4f m;
    for (int i = 0; i < 16; ++i) m[i] = m16[i];
    const Mat4f inv = m.invert();
    for (int i = 0; i < 16; ++i) out16[i] = inv[i];
}
void ref_mat4f_transform_vector(const float m16[16], const float v[3], float out[3]) {
    Mat4f m;
    for (int i = 0; i < 16; ++i) m[i] = m16[i];
    const Vec3f r = m * Vec3f(v[0], v[1], v[2]);
    out[0] = r.x(); out[1] = r.y(); out[2] = r.z();
}

}  // extern "C"
