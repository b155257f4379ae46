// stinput.h - SolTrace STINPUT text files as a plain document model.
//
// The file format (data/stinput/toy_problem_parabolic.stinput; SolTrace's own
// "# SOLTRACE VERSION ... INPUT FILE" layout) is read into a Document first and
// only then applied to a SolTraceSystem (SolTraceSystem::read_st_input), so the
// parser knows nothing about elements, surfaces or the device.
//
// Layout, one record per line, fields separated by tabs:
//   # SOLTRACE VERSION a.b.c INPUT FILE ...
//   SUN  PTSRC p  SHAPE c  SIGMA s  HALFWIDTH h       (s, h in mrad)
//   XYZ x y z  USELDH u  LDH lat day hour
//   USER SHAPE DATA n            followed by n lines "angle intensity"
//   OPTICS LIST COUNT n          followed by n optical pairs:
//     OPTICAL PAIR <name>
//     OPTICAL <19 fields>        front side, then the back side; each may be
//                                followed by reflectivity / transmissivity tables
//   STAGE LIST COUNT n           followed by n stages:
//     STAGE XYZ x y z AIM x y z ZROT z VIRTUAL v MULTIHIT m ELEMENTS k TRACETHROUGH t
//     <stage name>
//     k element rows of 29 fields (ElementColumn)
#pragma once

#include <array>
#include <istream>
#include <string>
#include <vector>

namespace stinput {

// Column of each field in an element row.
enum ElementColumn : size_t {
    kEnabled = 0,
    kOrigin = 1,         // x y z
    kAim = 4,            // x y z
    kZrot = 7,           // degrees
    kApertureCode = 8,   // 'r' rectangle, 'l' single-axis curvature section, 'c' circle, ...
    kApertureParams = 9, // 8 numbers
    kSurfaceCode = 17,   // 'f' flat, 'p' parabolic, 't' cylindrical, ...
    kSurfaceParams = 18, // 8 numbers
    kSurfaceFile = 26,
    kOpticName = 27,
    kInteraction = 28,
    kElementColumns = 29
};

// Column of each field in an OPTICAL row.
enum OpticColumn : size_t {
    kDistribution = 1,     // 'g' Gaussian, 'p' pillbox
    kApertureStop = 2,
    kSurfaceNumber = 3,
    kDiffractionOrder = 4,
    kReflectivity = 5,
    kTransmissivity = 6,
    kRmsSlope = 7,         // mrad
    kRmsSpecularity = 8,   // mrad
    kRefractionReal = 9,
    kRefractionImag = 10,
    kGratingCoeffs = 11,   // 4 numbers
    kUseReflTable = 15,
    kReflPoints = 16,
    kUseTransTable = 17,
    kTransPoints = 18,
    kOpticMinColumns = 15
};

struct Sun {
    bool point_source = false;
    char shape = 'g';
    double sigma_mrad = 0, halfwidth_mrad = 0;
    std::array<double, 3> xyz{0, 0, 1};
    bool use_ldh = false;
    std::array<double, 3> ldh{0, 0, 0};  // latitude, day, hour
    std::vector<std::array<double, 2>> user_shape;
};

struct OpticSide {
    char distribution = 'g';
    double reflectivity = 1, transmissivity = 0, rms_slope_mrad = 0, rms_specularity_mrad = 0;
    std::vector<std::array<double, 2>> reflectivity_table, transmissivity_table;
};

struct OpticPair {
    std::string name;
    OpticSide front, back;
};

struct ElementRow {
    bool enabled = true;
    std::array<double, 3> origin{0, 0, 0}, aim{0, 0, 1};
    double zrot_deg = 0;
    char aperture = 0, surface = 0;
    std::array<double, 8> aperture_params{}, surface_params{};
    std::string surface_file, optic;
    int interaction = 0;
    int line = 0;  // 1-based line of the row in the file
};

struct Stage {
    std::string name;
    std::array<double, 3> origin{0, 0, 0}, aim{0, 0, 1};
    double zrot_deg = 0;
    bool is_virtual = false, multihit = true, tracethrough = false;
    std::vector<ElementRow> elements;
};

struct Document {
    std::array<int, 3> version{0, 0, 0};
    Sun sun;
    std::vector<OpticPair> optics;
    std::vector<Stage> stages;
    const OpticPair* find_optic(const std::string& name) const;
};

// Parses STINPUT text. On malformed input returns false with `error` naming the
// line and what was expected there; `doc` is then incomplete.
bool parse(std::istream& in, Document& doc, std::string& error);
bool parse_file(const std::string& path, Document& doc, std::string& error);

}  // namespace stinput
