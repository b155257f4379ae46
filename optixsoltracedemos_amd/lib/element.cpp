// element.cpp - FP64 element model (lib/element.cpp).
#include "element.h"

#include <algorithm>
#include <limits>
#include <stdexcept>

ApertureType Element::get_aperture_type() const {
    if (!aperture_) throw std::runtime_error("element has no aperture");
    return aperture_->get_aperture_type();
}

SurfaceType Element::get_surface_type() const {
    if (!surface_) throw std::runtime_error("element has no surface");
    return surface_->get_surface_type();
}

void Element::update_euler_angles(const Vector3d& aim_point, double zrot_deg) {
    euler_ = mathUtil::normal_to_euler((aim_point - origin_).normalized(), zrot_deg);
}

void Element::update_euler_angles() { update_euler_angles(aim_, zrot_); }

void Element::update_element(const Vector3d& aim_point, double zrot_deg) {
    aim_ = aim_point;
    zrot_ = zrot_deg;
    update_euler_angles();
}

Matrix33d Element::get_rotation_matrix() const { return mathUtil::get_rotation_matrix_G2L(euler_).transpose(); }

void Element::compute_bounding_box() {
    const Matrix33d R = get_rotation_matrix();
    const double w = aperture_ ? aperture_->get_width() : 0.0, h = aperture_ ? aperture_->get_height() : 0.0;
    const bool cylinder = surface_ && surface_->get_surface_type() == SurfaceType::CYLINDER;
    if (!cylinder && (!aperture_ || aperture_->get_aperture_type() != ApertureType::RECTANGLE)) return;
    const int corners = cylinder ? 8 : 4;  // cylinder: (+-w/2, +-h/2, +-w/2) box
    double lo[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    std::numeric_limits<double>::max()};
    double hi[3] = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
                    std::numeric_limits<double>::lowest()};
    for (int c = 0; c < corners; ++c) {
        const Vector3d local((c & 1) ? w / 2 : -w / 2, (c & 2) ? h / 2 : -h / 2,
                             cylinder ? ((c & 4) ? w / 2 : -w / 2) : 0.0);
        const Vector3d g = R * local + origin_;
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], g[k]);
            hi[k] = std::max(hi[k], g[k]);
        }
    }
    lower_ = Vector3d(lo[0], lo[1], lo[2]);
    upper_ = Vector3d(hi[0], hi[1], hi[2]);
}

GeometryDataST Element::toDeviceGeometryData() const {
    GeometryDataST g;
    if (!aperture_ || !surface_ || aperture_->get_aperture_type() != ApertureType::RECTANGLE) return g;
    const double w = aperture_->get_width(), h = aperture_->get_height();
    const Matrix33d R = get_rotation_matrix();
    const Vector3d xb = R.get_x_basis(), yb = R.get_y_basis(), zb = R.get_z_basis();
    switch (surface_->get_surface_type()) {
        case SurfaceType::FLAT:
            g.setRectangle_Flat(GeometryDataST::Rectangle_Flat(mathUtil::toFloat3(origin_), mathUtil::toFloat3(xb),
                                                               mathUtil::toFloat3(yb), (float)w, (float)h));
            break;
        case SurfaceType::PARABOLIC: {
            // v1 = -w x_b, v2 = h y_b, anchor = origin - v1/2 - v2/2 (lib/element.cpp:138-146)
            const Vector3d v1 = xb * (-w), v2 = yb * h;
            g.setRectangleParabolic(GeometryDataST::Rectangle_Parabolic(
                mathUtil::toFloat3(v1), mathUtil::toFloat3(v2), mathUtil::toFloat3(origin_ - v1 * 0.5 - v2 * 0.5),
                (float)surface_->get_curvature_1(), (float)surface_->get_curvature_2()));
            break;
        }
        case SurfaceType::CYLINDER:
            g.setCylinder_Y(GeometryDataST::Cylinder_Y(mathUtil::toFloat3(origin_), (float)(w / 2), (float)(h / 2),
                                                       mathUtil::toFloat3(xb), mathUtil::toFloat3(zb)));
            break;
        default:
            break;
    }
    return g;
}
