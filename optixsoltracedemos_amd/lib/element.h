// element.h - optical element of the drop-in API (lib/element.h:41-102), plus
// per-element optics (reflectivity, slope and specular error) that the
// reference parses from STINPUT but never applies.
#ifndef SOLTRACE_ELEMENT_H
#define SOLTRACE_ELEMENT_H

#include <memory>

#include "Aperture.h"
#include "GeometryDataST.h"
#include "Surface.h"
#include "math_util.h"
#include "soltrace_type.h"
#include "vector3d.h"

class ElementBase {
public:
    ElementBase() = default;
    virtual ~ElementBase() = default;
    virtual const Vector3d& get_origin() const = 0;
    virtual void set_origin(const Vector3d&) = 0;
    virtual const Vector3d& get_aim_point() const = 0;
    virtual void set_aim_point(const Vector3d&) = 0;
    virtual GeometryDataST toDeviceGeometryData() const = 0;
};

class Element : public ElementBase {
public:
    Element() = default;
    ~Element() override = default;

    const Vector3d& get_origin() const override { return origin_; }
    void set_origin(const Vector3d& o) override { origin_ = o; }
    void set_aim_point(const Vector3d& a) override { aim_ = a; }
    const Vector3d& get_aim_point() const override { return aim_; }
    void set_zrot(double zrot_deg) { zrot_ = zrot_deg; }
    double get_zrot() const { return zrot_; }

    std::shared_ptr<Aperture> get_aperture() const { return aperture_; }
    std::shared_ptr<Surface> get_surface() const { return surface_; }
    ApertureType get_aperture_type() const;
    SurfaceType get_surface_type() const;
    void set_aperture(const std::shared_ptr<Aperture>& a) { aperture_ = a; }
    void set_surface(const std::shared_ptr<Surface>& s) { surface_ = s; }

    // orientation from an aim point and zrot (degrees), lib/element.cpp:77-93
    void update_euler_angles(const Vector3d& aim_point, double zrot_deg);
    void update_euler_angles();
    void update_element(const Vector3d& aim_point, double zrot_deg);
    const Vector3d& get_euler_angles() const { return euler_; }

    // local-to-global rotation (transpose of Spencer-Murty G2L)
    Matrix33d get_rotation_matrix() const;

    // reference AABB (lib/element.cpp:172-275)
    void compute_bounding_box();
    Vector3d get_upper_bounding_box() const { return upper_; }
    Vector3d get_lower_bounding_box() const { return lower_; }

    GeometryDataST toDeviceGeometryData() const override;

    // optics (new): defaults reflect every ray with no error
    void set_reflectivity(double r) { reflectivity_ = r; }
    double get_reflectivity() const { return reflectivity_; }
    void set_slope_error(double rad) { slope_error_ = rad; }
    double get_slope_error() const { return slope_error_; }
    void set_specularity_error(double rad) { spec_error_ = rad; }
    double get_specularity_error() const { return spec_error_; }

private:
    Vector3d origin_{0.0, 0.0, 0.0};
    Vector3d aim_{0.0, 0.0, 1.0};
    Vector3d euler_{0.0, 0.0, 0.0};
    double zrot_ = 0.0;
    Vector3d upper_, lower_;
    std::shared_ptr<Surface> surface_;
    std::shared_ptr<Aperture> aperture_;
    double reflectivity_ = 1.0, slope_error_ = 0.0, spec_error_ = 0.0;
};

#endif
