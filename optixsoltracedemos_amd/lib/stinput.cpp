// stinput.cpp - STINPUT text -> stinput::Document (format described in stinput.h).
#include "stinput.h"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace stinput {
namespace {

std::vector<std::string> tab_fields(const std::string& line) {
    std::vector<std::string> f;
    size_t start = 0;
    for (;;) {
        const size_t tab = line.find('\t', start);
        f.push_back(line.substr(start, tab == std::string::npos ? std::string::npos : tab - start));
        if (tab == std::string::npos) break;
        start = tab + 1;
    }
    return f;
}

std::string trim(const std::string& s) {
    const size_t b = s.find_first_not_of(" \t");
    if (b == std::string::npos) return std::string();
    const size_t e = s.find_last_not_of(" \t");
    return s.substr(b, e - b + 1);
}

// The file, line by line, with the position of the next unread line.
class Lines {
public:
    explicit Lines(std::istream& in) {
        std::string l;
        while (std::getline(in, l)) {
            while (!l.empty() && (l.back() == '\r' || l.back() == '\n')) l.pop_back();
            text_.push_back(l);
        }
    }
    bool done() const { return next_ >= text_.size(); }
    int number() const { return (int)next_; }  // 1-based number of the line last taken
    const std::string& take() {
        static const std::string empty;
        if (done()) {
            ++next_;
            return empty;
        }
        return text_[next_++];
    }

private:
    std::vector<std::string> text_;
    size_t next_ = 0;
};

struct Failure {
    std::string what;
};

[[noreturn]] void fail_at(const Lines& ln, const std::string& what) {
    throw Failure{"line " + std::to_string(ln.number()) + ": " + what};
}

// A number field; an empty field reads as 0, as SolTrace writes blank columns.
double number(const Lines& ln, const std::string& field, const char* name) {
    const std::string t = trim(field);
    if (t.empty()) return 0.0;
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != 0 || errno == ERANGE) fail_at(ln, std::string(name) + " is not a number: '" + t + "'");
    return v;
}

int integer(const Lines& ln, const std::string& field, const char* name) {
    const double v = number(ln, field, name);
    if (v != (double)(int)v) fail_at(ln, std::string(name) + " is not an integer: '" + trim(field) + "'");
    return (int)v;
}

// Lines of the form TAG key value [value...] key value ...: the values that
// follow `key`.
struct Tagged {
    const Lines& ln;
    std::vector<std::string> f;
    Tagged(const Lines& l, const std::string& text) : ln(l), f(tab_fields(text)) {}
    size_t at(const char* key, size_t values) const {
        for (size_t i = 0; i < f.size(); ++i)
            if (trim(f[i]) == key) {
                if (i + values >= f.size()) fail_at(ln, std::string("too few values after ") + key);
                return i + 1;
            }
        fail_at(ln, std::string("no ") + key + " field in '" + f[0] + "' record");
    }
    bool has(const char* key) const {
        for (const auto& x : f)
            if (trim(x) == key) return true;
        return false;
    }
    double num(const char* key, size_t k = 0) const { return number(ln, f[at(key, k + 1) + k], key); }
};

// "TAG<tab>n" count records (USER SHAPE DATA, OPTICS LIST COUNT, STAGE LIST COUNT)
int count_record(Lines& ln, const char* tag) {
    const std::vector<std::string> f = tab_fields(ln.take());
    if (trim(f[0]) != tag || f.size() < 2) fail_at(ln, std::string("expected '") + tag + "<tab>count'");
    const int n = integer(ln, f[1], tag);
    if (n < 0) fail_at(ln, std::string("negative ") + tag);
    return n;
}

// rows of two whitespace-separated numbers (sun shape and optic tables)
void read_pairs(Lines& ln, int n, std::vector<std::array<double, 2>>& out, const char* what) {
    out.clear();
    for (int i = 0; i < n; ++i) {
        std::istringstream row(ln.take());
        std::array<double, 2> p{};
        if (!(row >> p[0] >> p[1])) fail_at(ln, std::string("expected two numbers in the ") + what + " table");
        out.push_back(p);
    }
}

void read_header(Lines& ln, Document& doc) {
    const std::string& first = ln.take();
    if (first.empty() || first[0] != '#') fail_at(ln, "not a STINPUT file (the first line must begin with '#')");
    const size_t v = first.find("VERSION");
    if (v != std::string::npos) {
        std::istringstream s(first.substr(v + 7));
        char dot = 0;
        s >> doc.version[0] >> dot >> doc.version[1] >> dot >> doc.version[2];
    }
}

void read_sun(Lines& ln, Sun& sun) {
    {
        Tagged t(ln, ln.take());
        if (trim(t.f[0]) != "SUN") fail_at(ln, "expected the SUN record");
        sun.point_source = t.num("PTSRC") != 0;
        const std::string shape = trim(t.f[t.at("SHAPE", 1)]);
        sun.shape = shape.empty() ? 'g' : (char)std::tolower((unsigned char)shape[0]);
        sun.sigma_mrad = t.num("SIGMA");
        sun.halfwidth_mrad = t.has("HALFWIDTH") ? t.num("HALFWIDTH") : sun.sigma_mrad;
    }
    {
        Tagged t(ln, ln.take());
        if (trim(t.f[0]) != "XYZ" || t.f.size() < 4) fail_at(ln, "expected the sun's XYZ record");
        for (int k = 0; k < 3; ++k) sun.xyz[k] = number(ln, t.f[1 + k], "sun XYZ");
        sun.use_ldh = t.has("USELDH") && t.num("USELDH") != 0;
        if (t.has("LDH"))
            for (int k = 0; k < 3; ++k) sun.ldh[k] = t.num("LDH", k);
    }
    read_pairs(ln, count_record(ln, "USER SHAPE DATA"), sun.user_shape, "user sun shape");
}

void read_optic_side(Lines& ln, OpticSide& side) {
    const std::vector<std::string> f = tab_fields(ln.take());
    if (trim(f[0]) != "OPTICAL") fail_at(ln, "expected an OPTICAL record");
    if (f.size() < kOpticMinColumns)
        fail_at(ln, "OPTICAL record has " + std::to_string(f.size()) + " fields, at least " +
                        std::to_string((int)kOpticMinColumns) + " expected");
    const std::string dist = trim(f[kDistribution]);
    side.distribution = dist.empty() ? 'g' : dist[0];
    side.reflectivity = number(ln, f[kReflectivity], "reflectivity");
    side.transmissivity = number(ln, f[kTransmissivity], "transmissivity");
    side.rms_slope_mrad = number(ln, f[kRmsSlope], "RMS slope");
    side.rms_specularity_mrad = number(ln, f[kRmsSpecularity], "RMS specularity");
    const bool refl = f.size() > kReflPoints && integer(ln, f[kUseReflTable], "reflectivity table flag") > 0;
    const int refl_n = f.size() > kReflPoints ? integer(ln, f[kReflPoints], "reflectivity table size") : 0;
    const bool trans = f.size() > kTransPoints && integer(ln, f[kUseTransTable], "transmissivity table flag") > 0;
    const int trans_n = f.size() > kTransPoints ? integer(ln, f[kTransPoints], "transmissivity table size") : 0;
    if (refl) read_pairs(ln, refl_n, side.reflectivity_table, "reflectivity");
    if (trans) read_pairs(ln, trans_n, side.transmissivity_table, "transmissivity");
}

void read_optics(Lines& ln, std::vector<OpticPair>& optics) {
    const int n = count_record(ln, "OPTICS LIST COUNT");
    for (int i = 0; i < n; ++i) {
        const std::vector<std::string> f = tab_fields(ln.take());
        if (trim(f[0]).rfind("OPTICAL PAIR", 0) != 0) fail_at(ln, "expected 'OPTICAL PAIR<tab>name'");
        OpticPair p;
        p.name = f.size() > 1 ? trim(f[1]) : trim(trim(f[0]).substr(12));
        read_optic_side(ln, p.front);
        read_optic_side(ln, p.back);
        optics.push_back(std::move(p));
    }
}

void read_element_row(Lines& ln, ElementRow& e) {
    const std::vector<std::string> f = tab_fields(ln.take());
    e.line = ln.number();
    if (f.size() < kElementColumns)
        fail_at(ln, "element row has " + std::to_string(f.size()) + " fields, " +
                        std::to_string((int)kElementColumns) + " expected");
    e.enabled = integer(ln, f[kEnabled], "enabled flag") != 0;
    for (int k = 0; k < 3; ++k) {
        e.origin[k] = number(ln, f[kOrigin + k], "origin");
        e.aim[k] = number(ln, f[kAim + k], "aim point");
    }
    e.zrot_deg = number(ln, f[kZrot], "z rotation");
    const std::string ap = trim(f[kApertureCode]), sf = trim(f[kSurfaceCode]);
    if (ap.empty() || sf.empty()) fail_at(ln, "element row without an aperture or surface code");
    e.aperture = ap[0];
    e.surface = sf[0];
    for (int k = 0; k < 8; ++k) {
        e.aperture_params[k] = number(ln, f[kApertureParams + k], "aperture parameter");
        e.surface_params[k] = number(ln, f[kSurfaceParams + k], "surface parameter");
    }
    e.surface_file = trim(f[kSurfaceFile]);
    e.optic = trim(f[kOpticName]);
    e.interaction = integer(ln, f[kInteraction], "interaction type");
}

void read_stages(Lines& ln, std::vector<Stage>& stages) {
    const int n = count_record(ln, "STAGE LIST COUNT");
    for (int i = 0; i < n; ++i) {
        Stage st;
        int rows = 0;
        {
            Tagged t(ln, ln.take());
            if (trim(t.f[0]) != "STAGE") fail_at(ln, "expected a STAGE record");
            for (int k = 0; k < 3; ++k) {
                st.origin[k] = t.num("XYZ", k);
                st.aim[k] = t.num("AIM", k);
            }
            st.zrot_deg = t.num("ZROT");
            st.is_virtual = t.num("VIRTUAL") != 0;
            st.multihit = t.num("MULTIHIT") != 0;
            rows = (int)t.num("ELEMENTS");
            st.tracethrough = t.has("TRACETHROUGH") && t.num("TRACETHROUGH") != 0;
            if (rows < 0) fail_at(ln, "negative element count");
        }
        st.name = ln.take();
        st.elements.resize((size_t)rows);
        for (ElementRow& e : st.elements) read_element_row(ln, e);
        stages.push_back(std::move(st));
    }
}

}  // namespace

const OpticPair* Document::find_optic(const std::string& name) const {
    for (const OpticPair& p : optics)
        if (p.name == name) return &p;
    return nullptr;
}

bool parse(std::istream& in, Document& doc, std::string& error) {
    Lines ln(in);
    try {
        read_header(ln, doc);
        read_sun(ln, doc.sun);
        read_optics(ln, doc.optics);
        read_stages(ln, doc.stages);
    } catch (const Failure& f) {
        error = f.what;
        return false;
    }
    return true;
}

bool parse_file(const std::string& path, Document& doc, std::string& error) {
    std::ifstream in(path);
    if (!in) {
        error = "cannot open '" + path + "'";
        return false;
    }
    return parse(in, doc, error);
}

}  // namespace stinput
