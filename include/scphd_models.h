/*
 * scphd_models.h -- motion and measurement model plugins of the SC-PHD
 * calibration filter, usable in host code and in HIP device code.
 *
 * Names, constructor arguments and method meanings follow the reference
 * plugin surface (cheesinglee/scphd-camera-calibration):
 *   EuclideanPoint, DisparityPoint, Extrinsics   src/types.h:55-140
 *   LinearCVMotionModel3D                        src/linearcvmotionmodel3d.cuh:12-50
 *   OrientedLinearCVMotionModel3D                src/OrientedLinearCVMotionModel3D.cuh:13-67
 *   DisparityMeasurementModel                    src/disparitymeasurementmodel.cuh:15-51
 *   transformWorldToCamera / transformCameraToWorld  src/disparitymeasurementmodel.cu:3-144
 *   wrapAngle                                    src/device_math.cu:308-317
 * All types are trivially copyable and are passed by value into kernels.
 *
 * Additions with no reference counterpart: CameraPose (a rotation and
 * translation evaluated once per parent so per-sample transforms need no
 * trigonometry) and DisparityMeasurementModel::inRange (the field-of-view
 * test the reference writes inline in worldToDisparityKernel,
 * src/kernels.cu:37-45).
 */
#ifndef SCPHD_MODELS_H
#define SCPHD_MODELS_H

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SCPHD_HD __host__ __device__
#else
#define SCPHD_HD
#endif

struct EuclideanPoint {
    double x, y, z, vx, vy, vz;
};

struct DisparityPoint {
    double u, v, d, vu, vv, vd;
};

struct Extrinsics {
    EuclideanPoint cartesian;
    EuclideanPoint angular;
};

/* Wrap an angle to [-pi, pi] (fmod by 2 pi, then one shift). */
SCPHD_HD inline double wrapAngle(double a) {
    double r = fmod(a, 2.0 * M_PI);
    if (r > M_PI)
        r -= 2.0 * M_PI;
    else if (r < -M_PI)
        r += 2.0 * M_PI;
    return r;
}

/* Rigid camera pose: p_world = R p_camera + t, with
 * R = Rx(angular.x) Ry(angular.y) Rz(angular.z), row-major. */
struct CameraPose {
    double R[9];
    double t[3];

    SCPHD_HD static CameraPose identity() {
        CameraPose c;
        for (int i = 0; i < 9; ++i) c.R[i] = (i % 4 == 0) ? 1.0 : 0.0;
        c.t[0] = c.t[1] = c.t[2] = 0.0;
        return c;
    }

    SCPHD_HD static CameraPose fromExtrinsics(const Extrinsics& e) {
        const double ct = cos(e.angular.x), st = sin(e.angular.x);
        const double cp = cos(e.angular.y), sp = sin(e.angular.y);
        const double cs = cos(e.angular.z), ss = sin(e.angular.z);
        CameraPose c;
        c.R[0] = cp * cs;                c.R[1] = -cp * ss;               c.R[2] = sp;
        c.R[3] = ct * ss + st * sp * cs; c.R[4] = ct * cs - st * sp * ss; c.R[5] = -st * cp;
        c.R[6] = st * ss - ct * sp * cs; c.R[7] = st * cs + ct * sp * ss; c.R[8] = ct * cp;
        c.t[0] = e.cartesian.x;
        c.t[1] = e.cartesian.y;
        c.t[2] = e.cartesian.z;
        return c;
    }

    /* p_camera = R^T (p_world - t), velocities rotated only. */
    SCPHD_HD EuclideanPoint worldToCamera(const EuclideanPoint& p) const {
        const double dx = p.x - t[0], dy = p.y - t[1], dz = p.z - t[2];
        EuclideanPoint c;
        c.x = R[0] * dx + R[3] * dy + R[6] * dz;
        c.y = R[1] * dx + R[4] * dy + R[7] * dz;
        c.z = R[2] * dx + R[5] * dy + R[8] * dz;
        c.vx = R[0] * p.vx + R[3] * p.vy + R[6] * p.vz;
        c.vy = R[1] * p.vx + R[4] * p.vy + R[7] * p.vz;
        c.vz = R[2] * p.vx + R[5] * p.vy + R[8] * p.vz;
        return c;
    }

    /* p_world = R p_camera + t, velocities rotated only. */
    SCPHD_HD EuclideanPoint cameraToWorld(const EuclideanPoint& p) const {
        EuclideanPoint w;
        w.x = R[0] * p.x + R[1] * p.y + R[2] * p.z + t[0];
        w.y = R[3] * p.x + R[4] * p.y + R[5] * p.z + t[1];
        w.z = R[6] * p.x + R[7] * p.y + R[8] * p.z + t[2];
        w.vx = R[0] * p.vx + R[1] * p.vy + R[2] * p.vz;
        w.vy = R[3] * p.vx + R[4] * p.vy + R[5] * p.vz;
        w.vz = R[6] * p.vx + R[7] * p.vy + R[8] * p.vz;
        return w;
    }
};

SCPHD_HD inline EuclideanPoint transformWorldToCamera(EuclideanPoint p_world, Extrinsics e) {
    return CameraPose::fromExtrinsics(e).worldToCamera(p_world);
}

SCPHD_HD inline EuclideanPoint transformCameraToWorld(EuclideanPoint p_camera, Extrinsics e) {
    return CameraPose::fromExtrinsics(e).cameraToWorld(p_camera);
}

/* Constant-velocity motion with acceleration noise per axis. */
class LinearCVMotionModel3D {
public:
    SCPHD_HD LinearCVMotionModel3D() : std_ax_(1.0), std_ay_(1.0), std_az_(1.0) {}
    SCPHD_HD LinearCVMotionModel3D(double ax, double ay, double az) : std_ax_(ax), std_ay_(ay), std_az_(az) {}

    SCPHD_HD EuclideanPoint computeMotion(EuclideanPoint s, double dt) const {
        s.x = s.x + s.vx * dt;
        s.y = s.y + s.vy * dt;
        s.z = s.z + s.vz * dt;
        return s;
    }

    /* ax, ay, az are acceleration draws (already scaled by the std). */
    SCPHD_HD EuclideanPoint computeNoisyMotion(EuclideanPoint s, double dt, double ax, double ay,
                                               double az) const {
        EuclideanPoint n = computeMotion(s, dt);
        n.x += 0.5 * ax * dt * dt;
        n.y += 0.5 * ay * dt * dt;
        n.z += 0.5 * az * dt * dt;
        n.vx += ax * dt;
        n.vy += ay * dt;
        n.vz += az * dt;
        return n;
    }

    SCPHD_HD double std_ax() const { return std_ax_; }
    SCPHD_HD double std_ay() const { return std_ay_; }
    SCPHD_HD double std_az() const { return std_az_; }

private:
    double std_ax_, std_ay_, std_az_;
};

/* A pair of CV models: cartesian and angular parts of the extrinsics; the
 * angles are wrapped after every step. */
class OrientedLinearCVMotionModel3D {
public:
    SCPHD_HD OrientedLinearCVMotionModel3D() {}
    SCPHD_HD OrientedLinearCVMotionModel3D(double ax, double ay, double az, double ax_angular,
                                           double ay_angular, double az_angular)
        : cartesian_model_(ax, ay, az), angular_model_(ax_angular, ay_angular, az_angular) {}

    SCPHD_HD void computeMotion(EuclideanPoint& cartesian_state, EuclideanPoint& angular_state,
                                double dt) const {
        cartesian_state = cartesian_model_.computeMotion(cartesian_state, dt);
        angular_state = angular_model_.computeMotion(angular_state, dt);
        angular_state.x = wrapAngle(angular_state.x);
        angular_state.y = wrapAngle(angular_state.y);
        angular_state.z = wrapAngle(angular_state.z);
    }

    SCPHD_HD void computeNoisyMotion(EuclideanPoint& cartesian_state, EuclideanPoint& angular_state,
                                     double dt, double ax_c, double ay_c, double az_c, double ax_a,
                                     double ay_a, double az_a) const {
        cartesian_state = cartesian_model_.computeNoisyMotion(cartesian_state, dt, ax_c, ay_c, az_c);
        angular_state = angular_model_.computeNoisyMotion(angular_state, dt, ax_a, ay_a, az_a);
        angular_state.x = wrapAngle(angular_state.x);
        angular_state.y = wrapAngle(angular_state.y);
        angular_state.z = wrapAngle(angular_state.z);
    }

    SCPHD_HD double std_ax_cartesian() const { return cartesian_model_.std_ax(); }
    SCPHD_HD double std_ay_cartesian() const { return cartesian_model_.std_ay(); }
    SCPHD_HD double std_az_cartesian() const { return cartesian_model_.std_az(); }
    SCPHD_HD double std_ax_angular() const { return angular_model_.std_ax(); }
    SCPHD_HD double std_ay_angular() const { return angular_model_.std_ay(); }
    SCPHD_HD double std_az_angular() const { return angular_model_.std_az(); }

private:
    LinearCVMotionModel3D cartesian_model_;
    LinearCVMotionModel3D angular_model_;
};

/* Pinhole camera in "disparity space" (u, v, d = -fx/z and their rates);
 * image = 2 u0 x 2 v0, clutter density kappa = lambda / (2 v0)^2. */
class DisparityMeasurementModel {
public:
    SCPHD_HD DisparityMeasurementModel()
        : fx_(1000), fy_(1000), u0_(250), v0_(250), img_width_(500), img_height_(500), std_u_(1),
          std_v_(1), pd_(0.99), kappa_(0) {}

    SCPHD_HD DisparityMeasurementModel(double fx, double fy, double u0, double v0, double std_u,
                                       double std_v, double pd, double lambda)
        : fx_(fx), fy_(fy), u0_(u0), v0_(v0), img_width_(2 * u0), img_height_(2 * v0), std_u_(std_u),
          std_v_(std_v), pd_(pd), kappa_(lambda / ((2 * v0) * (2 * v0))) {}

    /* Camera-frame point -> disparity point:
     *   u = u0 - fx x / z, v = v0 - fy y / z, d = -fx / z,
     *   u' = -fx (x' z - z' x) / z^2, v' = -fy (y' z - z' y) / z^2, d' = fx z' / z^2.
     * The z^2 quotients multiply by 1 / z^2 (one division instead of three). */
    SCPHD_HD DisparityPoint computeMeasurement(EuclideanPoint p) const {
        const double x = p.x, y = p.y, z = p.z;
        const double fz = fx_ / z, gz = fy_ / z, iz2 = 1.0 / (z * z);
        DisparityPoint q;
        q.u = u0_ - fz * x;
        q.v = v0_ - gz * y;
        q.d = -fz;
        q.vu = -fx_ * (p.vx * z - p.vz * x) * iz2;
        q.vv = -fy_ * (p.vy * z - p.vz * y) * iz2;
        q.vd = fx_ * p.vz * iz2;
        return q;
    }

    /* World point seen by a camera with extrinsics e. */
    SCPHD_HD DisparityPoint computeMeasurement(EuclideanPoint p_world, Extrinsics e) const {
        return computeMeasurement(transformWorldToCamera(p_world, e));
    }

    SCPHD_HD DisparityPoint computeMeasurement(EuclideanPoint p_world, const CameraPose& c) const {
        return computeMeasurement(c.worldToCamera(p_world));
    }

    /* Disparity point -> camera-frame point:
     *   x = (u - u0) / d, y = (fx/fy) (v - v0) / d, z = -fx / d,
     *   x' = (u' d - d' (u - u0)) / d^2, y' = (fx/fy) (v' d - d' (v - v0)) / d^2,
     *   z' = fx d' / d^2,
     * with the quotients taken as products by 1/d and 1/d^2. */
    SCPHD_HD EuclideanPoint invertToCamera(DisparityPoint q) const {
        const double u = q.u, v = q.v, d = q.d;
        const double id = 1.0 / d, id2 = 1.0 / (d * d), r = fx_ / fy_;
        EuclideanPoint c;
        c.x = (u - u0_) * id;
        c.y = r * (v - v0_) * id;
        c.z = -fx_ * id;
        c.vx = (q.vu * d - q.vd * (u - u0_)) * id2;
        c.vy = r * (q.vv * d - q.vd * (v - v0_)) * id2;
        c.vz = fx_ * q.vd * id2;
        return c;
    }

    SCPHD_HD EuclideanPoint invertMeasurement(DisparityPoint q, Extrinsics e) const {
        return transformCameraToWorld(invertToCamera(q), e);
    }

    SCPHD_HD EuclideanPoint invertMeasurement(DisparityPoint q, const CameraPose& c) const {
        return c.cameraToWorld(invertToCamera(q));
    }

    /* 0 <= u <= width, 0 <= v <= height, d >= 0 */
    SCPHD_HD bool inRange(const DisparityPoint& q) const {
        return q.u >= 0 && q.u <= img_width_ && q.v >= 0 && q.v <= img_height_ && q.d >= 0;
    }

    SCPHD_HD double std_u() const { return std_u_; }
    SCPHD_HD double std_v() const { return std_v_; }
    SCPHD_HD double pd() const { return pd_; }
    SCPHD_HD double kappa() const { return kappa_; }
    SCPHD_HD double img_width() const { return img_width_; }
    SCPHD_HD double img_height() const { return img_height_; }
    SCPHD_HD double fx() const { return fx_; }
    SCPHD_HD double fy() const { return fy_; }
    SCPHD_HD double u0() const { return u0_; }
    SCPHD_HD double v0() const { return v0_; }

private:
    double fx_, fy_, u0_, v0_, img_width_, img_height_, std_u_, std_v_, pd_, kappa_;
};

#endif /* SCPHD_MODELS_H */
