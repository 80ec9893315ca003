/*
 * kmc_color.cpp -- the surface colour of the reference, on the host.
 *
 * kvs::MarchingCubes colours the whole surface with one RGB value,
 * transferFunction().colorMap().at(isolevel) (MarchingCubes.cpp:239,286), after
 * MapperBase::setRange has given the transfer function the volume's value
 * range (MapperBase.cpp:54-83).  This file restates, in float arithmetic and
 * with the same operation order:
 *   - ColorMap::setPoints / create (ColorMap.cpp:879-895, 913-1019) for the
 *     BrewerRdBu (default, ColorMap.cpp:23,255-277) and BrewerSingleHueGreys
 *     (ColorMap.cpp:440-460) maps, interpolated in CIE Lab;
 *   - the sRGB <-> linear <-> XYZ <-> Lab conversions (Image/XYZColor.cpp:17-84,
 *     Image/LabColor.cpp:19-104) with std::pow as the reference calls it;
 *   - RGBColor::Mix, operator=(Vec3), toVec3 (Image/RGBColor.cpp:23-31,125-134,170-175);
 *   - ColorMap::at (ColorMap.cpp:1043-1054).
 * The colour is one value per surface; it never touches the device.
 */
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <list>
#include <utility>
#include <vector>

#include "kmc.h"
#include "kmc_guard.h"

namespace {

struct Rgb { uint8_t r, g, b; };
struct V3 { float x, y, z; };

/* Math::Round (Utility/Math.h:110-114) */
inline int round_half_away(float a) { return a >= 0 ? (int)(a + 0.5) : -(int)(0.5 - a); }
/* Math::Clamp */
inline float clampf(float a, float lo, float hi) { return a <= lo ? lo : (a <= hi ? a : hi); }
/* Math::Mix: x*(1-a) + y*a */
inline float mixf(float x, float y, float a) { return x * (1 - a) + y * a; }
/* Math::Equal on floats, compared against the double 1e-6 */
inline bool equalf(float a, float b)
{
    const float d = a - b;
    return (double)(d < 0 ? -d : d) < 1.e-6;
}

inline V3 to_vec3(Rgb c)
{
    return { (float)c.r / 255.0f, (float)c.g / 255.0f, (float)c.b / 255.0f };
}

inline Rgb from_vec3(V3 v)
{
    const float r = clampf(v.x, 0.0f, 1.0f);
    const float g = clampf(v.y, 0.0f, 1.0f);
    const float b = clampf(v.z, 0.0f, 1.0f);
    return { (uint8_t)round_half_away(r * 255.0f), (uint8_t)round_half_away(g * 255.0f),
             (uint8_t)round_half_away(b * 255.0f) };
}

/* Vector3::dot, accumulated from zero. */
inline float dot3(const float m[3], V3 v)
{
    float r = 0;
    r += m[0] * v.x;
    r += m[1] * v.y;
    r += m[2] * v.z;
    return r;
}

inline float to_linear(float c)
{
    float cl = 0;
    if (c <= 0.04045f) cl = c / 12.92f;
    else cl = std::pow((c + 0.055f) / 1.055f, 2.4f);
    return clampf(cl, 0.0f, 1.0f);
}

inline float to_standard(float c)
{
    float cs = 0;
    if (c <= 0.0031308f) cs = 12.92f * c;
    else cs = 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f;
    return clampf(cs, 0.0f, 1.0f);
}

/* D65 reference white. */
const float kXn = 0.95047f, kYn = 1.0f, kZn = 1.08883f;

inline float lab_f(float t)
{
    if (t > 0.008856f) return std::pow(t, 1.0f / 3.0f);
    return 7.787037f * t + 16.0f / 116.0f;
}

inline float lab_finv(float t)
{
    if (t > 0.008856f) return std::pow(t, 3.0f);
    return (t - 16.0f / 116.0f) / 7.787037f;
}

struct Lab { float l, a, b; };

Lab rgb_to_lab(Rgb c)
{
    static const float M[3][3] = {
        { 0.412391f, 0.357584f, 0.180481f },
        { 0.212639f, 0.715169f, 0.072192f },
        { 0.019331f, 0.119195f, 0.950532f } };
    const V3 s = to_vec3(c);
    const V3 lin = { to_linear(s.x), to_linear(s.y), to_linear(s.z) };
    const float X = dot3(M[0], lin), Y = dot3(M[1], lin), Z = dot3(M[2], lin);
    const float l = 116.0f * (lab_f(Y / kYn) - 16.0f / 116.0f);
    const float a = 500.0f * (lab_f(X / kXn) - lab_f(Y / kYn));
    const float b = 200.0f * (lab_f(Y / kYn) - lab_f(Z / kZn));
    return { l, a, b };
}

Rgb lab_to_rgb(Lab c)
{
    static const float Minv[3][3] = {
        {  3.240970f, -1.537383f, -0.498611f },
        { -0.969244f,  1.875968f,  0.041555f },
        {  0.055630f, -0.203977f,  1.056972f } };
    const float y = (c.l + 16.0f) / 116.0f;
    const float x = c.a / 500.0f + y;
    const float z = y - c.b / 200.0f;
    const V3 xyz = { kXn * lab_finv(x), kYn * lab_finv(y), kZn * lab_finv(z) };
    const V3 lin = { dot3(Minv[0], xyz), dot3(Minv[1], xyz), dot3(Minv[2], xyz) };
    return from_vec3({ to_standard(lin.x), to_standard(lin.y), to_standard(lin.z) });
}

Rgb mix_rgb(Rgb c0, Rgb c1, float t)
{
    const V3 a = to_vec3(c0), b = to_vec3(c1);
    return from_vec3({ mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t) });
}

/* Control colours of the maps, in the order setPoints receives them. */
std::vector<Rgb> control_colors(int map)
{
    if (map == KMC_COLORMAP_BREWER_SINGLE_HUE_GREYS) {
        return { {255, 255, 255}, {240, 240, 240}, {217, 217, 217}, {189, 189, 189}, {150, 150, 150},
                 {115, 115, 115}, {82, 82, 82}, {37, 37, 37}, {0, 0, 0} };
    }
    /* BrewerRdBu, listed red -> blue and reversed before use (ColorMap.cpp:257-270) */
    std::vector<Rgb> c = { {103, 0, 31}, {178, 24, 43}, {214, 96, 77}, {244, 165, 130}, {253, 219, 199},
                           {247, 247, 247}, {209, 229, 240}, {146, 197, 222}, {67, 147, 195},
                           {33, 102, 172}, {5, 48, 97} };
    return std::vector<Rgb>(c.rbegin(), c.rend());
}

/* ColorMap(resolution) + setPoints(colors) + Lab colour space + create(). */
bool create_table(int map, uint32_t resolution, std::vector<uint8_t>& table)
{
    if (map != KMC_COLORMAP_BREWER_RDBU && map != KMC_COLORMAP_BREWER_SINGLE_HUE_GREYS) return false;
    if (resolution < 2) return false;
    const std::vector<Rgb> colors = control_colors(map);
    typedef std::pair<float, Rgb> Point;
    std::list<Point> points;
    const size_t ncolors = colors.size();
    const float stride = 1.0f / (ncolors - 1);
    points.push_back(Point(0.0f, colors.front()));
    for (size_t index = 1; index + 1 < ncolors; ++index) {
        const float value = (float)round_half_away((float)resolution * stride * (float)index);
        points.push_back(Point(value, colors[index]));
    }
    points.push_back(Point(float(resolution - 1), colors.back()));

    const float min_value = 0.0f;
    const float max_value = (float)(resolution - 1);
    points.sort([](const Point& p1, const Point& p2) { return p1.first < p2.first; });
    const Rgb black = {0, 0, 0}, white = {255, 255, 255};
    if (points.front().first > min_value) points.push_back(Point(min_value, black));
    if (points.back().first < max_value) points.push_back(Point(max_value, white));

    table.assign(3 * (size_t)resolution, 0);
    const float fstride = (max_value - min_value) / (float)(resolution - 1);
    float f = min_value;
    for (uint32_t i = 0; i < resolution; ++i, f += fstride) {
        auto p = points.begin();
        auto last = points.end();
        Rgb color = {0, 0, 0};
        Point p0(min_value, black);
        Point p1(max_value, white);
        while (p != last) {
            const float s = p->first;
            if (equalf(f, s)) { color = p->second; break; }
            else if (f < s) {
                p1 = *p;
                const float s0 = p0.first, s1 = p1.first;
                const float ratio = (f - s0) / (s1 - s0);
                const Lab c0 = rgb_to_lab(p0.second), c1 = rgb_to_lab(p1.second);
                const Lab m = { mixf(c0.l, c1.l, ratio), mixf(c0.a, c1.a, ratio), mixf(c0.b, c1.b, ratio) };
                color = lab_to_rgb(m);
                break;
            } else {
                p0 = *p;
                ++p;
                if (p == last && equalf(p0.first, max_value)) color = p0.second;
            }
        }
        table[3 * i + 0] = color.r;
        table[3 * i + 1] = color.g;
        table[3 * i + 2] = color.b;
    }
    return true;
}

}  // namespace

extern "C" {

kmc_status kmc_colormap_table(int32_t colormap, uint32_t resolution, uint8_t* rgb_out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!rgb_out) return KMC_ERR_INVALID_ARGUMENT;
        std::vector<uint8_t> t;
        if (!create_table(colormap, resolution, t)) return KMC_ERR_INVALID_ARGUMENT;
        for (size_t i = 0; i < t.size(); ++i) rgb_out[i] = t[i];
        return KMC_OK;
    });
}

kmc_status kmc_surface_color(int32_t colormap, uint32_t resolution, int32_t type, double volume_min,
                             double volume_max, double isolevel, uint8_t rgb[3])
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!rgb) return KMC_ERR_INVALID_ARGUMENT;
        std::vector<uint8_t> t;
        if (!create_table(colormap, resolution, t)) return KMC_ERR_INVALID_ARGUMENT;
        /* MapperBase::setRange (MapperBase.cpp:54-83) on a transfer function
         * without a range of its own. */
        float lo, hi;
        if (type == KMC_INT8) { lo = -128; hi = 127; }
        else if (type == KMC_UINT8) { lo = 0; hi = 255; }
        else {
            lo = (float)volume_min;
            hi = (float)volume_max;
            if (equalf(lo, hi)) hi = hi + 1.0f;
        }
        /* ColorMap::at(float(isolevel)) (ColorMap.cpp:1043-1054) */
        const float value = (float)isolevel;
        const float v0 = clampf(value, lo, hi);
        const float r = (float)(resolution - 1);
        float v = (v0 - lo) / (hi - lo) * r;
        /* when float(min) == float(max) and the +1 of setRange is absorbed
         * (|min| >= 2^24), hi - lo is 0 and v is NaN: the reference converts NaN to
         * size_t and reads outside its table; here the table's first colour */
        if (!(v >= 0.0f)) v = 0.0f;
        const size_t s0 = (size_t)v;
        const size_t s1 = (s0 + 1 < (size_t)(resolution - 1)) ? s0 + 1 : (size_t)(resolution - 1);
        const Rgb c0 = { t[3 * s0], t[3 * s0 + 1], t[3 * s0 + 2] };
        const Rgb c1 = { t[3 * s1], t[3 * s1 + 1], t[3 * s1 + 2] };
        const Rgb c = mix_rgb(c0, c1, v - (float)s0);
        rgb[0] = c.r; rgb[1] = c.g; rgb[2] = c.b;
        return KMC_OK;
    });
}

}  // extern "C"
