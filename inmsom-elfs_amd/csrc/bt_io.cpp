// DIOGIN direct-access records and GrADS descriptors: the reference's
// on-disk formats (Service/input_output_data.f90, Service/rw_ctl_file.f90).
//
// A DIOGIN file is a sequence of unformatted direct-access records with
// no markers: record r (1-based) starts at byte (r - 1) * recl, recl =
// (nxe - nxb + 1) * (nye - nyb + 1) * element size, and holds one level of
// the subdomain [nxb, nxe] x [nyb, nye] with the first index fastest.
// Field nfield, level kr (1-based within nzb..nze) is record
// (nfield - 1) * (nze - nzb + 1) + kr (input_output_data.f90:159,261).
// Land points (|lu| < 0.5) are written as undef = -1e32 and read back as 0.
//
// The descriptor writer reproduces ctl_file_write's text: its formats
// (Fortran E, G and I editing) are restated here with the Fortran rules, so
// a file this library writes is the one the reference's writer produces.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <fcntl.h>
#include <unistd.h>

#include "bt_internal.h"
#include "bt_io.h"
#include "inmsom_bt.h"

namespace ibt {
namespace io {
namespace {

constexpr float UNDEF = -1.0e32f;  // Inc/reclen.fi: undef

int fail(const std::string &m, int code)
{
    set_error(m);
    return code;
}

// fulfname (input_output_data.f90:4-114): path and file name without their
// blanks, joined by one '/'.
std::string fulfname(const char *path, const char *fname)
{
    auto strip = [](const char *s) {
        std::string t = s ? s : "";
        const size_t a = t.find_first_not_of(' ');
        if (a == std::string::npos) return std::string();
        const size_t b = t.find_last_not_of(' ');
        return t.substr(a, b - a + 1);
    };
    std::string p = strip(path), f = strip(fname);
    if (p.empty()) return f;
    const bool ps = p.back() == '/', fs = !f.empty() && f[0] == '/';
    if (ps && fs) p.pop_back();
    if (!ps && !fs) p.push_back('/');
    return p + f;
}

int check_box(int nx, int ny, int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze)
{
    if (nxe > nx || nxb < 1 || nxb > nxe) return fail("error in grid diapason of 1 - coordinate", INMSOM_BT_EINVAL);
    if (nye > ny || nyb < 1 || nyb > nye) return fail("error in grid diapason of 2 - coordinate", INMSOM_BT_EINVAL);
    if (nze > nz || nzb < 1 || nzb > nze) return fail("error in grid diapason of 3 - coordinate", INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

// wdstd / wdstd8: land points written as undef, then set to 0 in `field`
// (input_output_data.f90:268-300).
template <class T>
int write_records(const char *path, const char *fname, int nfield, T *field, const float *lu, int nx, int ny, int nz,
                  int nxb, int nxe, int nyb, int nye, int nzb, int nze)
{
    if (!field || !lu) return fail("null pointer", INMSOM_BT_EINVAL);
    if (nfield < 1) return fail("nfield must be >= 1", INMSOM_BT_EINVAL);
    int rc = check_box(nx, ny, nz, nxb, nxe, nyb, nye, nzb, nze);
    if (rc) return rc;
    const std::string name = fulfname(path, fname);
    const int fd = ::open(name.c_str(), O_WRONLY | O_CREAT, 0644);
    if (fd < 0) return fail("error in open file for writing: " + name + ": " + strerror(errno), INMSOM_BT_EINVAL);
    const int rx = nxe - nxb + 1, ry = nye - nyb + 1;
    const size_t recl = (size_t)rx * ry * sizeof(T);
    std::vector<T> rec((size_t)rx * ry);
    const long long nrecf = (long long)(nfield - 1) * (nze - nzb + 1);
    int kr = 0;
    for (int k = nzb; k <= nze; ++k) {
        ++kr;
        size_t q = 0;
        for (int j = nyb; j <= nye; ++j)
            for (int i = nxb; i <= nxe; ++i) {
                const size_t a = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1));
                const bool land = std::fabs(lu[(size_t)(i - 1) + (size_t)nx * (j - 1)]) < 0.5f;
                if (land) field[a] = (T)UNDEF;
                rec[q++] = field[a];
                if (land) field[a] = (T)0;
            }
        const off_t off = (off_t)(nrecf + kr - 1) * (off_t)recl;
        const char *b = (const char *)rec.data();
        size_t left = recl;
        while (left > 0) {
            const ssize_t w = ::pwrite(fd, b, left, off + (off_t)(recl - left));
            if (w <= 0) {
                ::close(fd);
                return fail("error in writing on file: " + name, INMSOM_BT_EINVAL);
            }
            b += w;
            left -= (size_t)w;
        }
    }
    if (::close(fd) != 0) return fail("error closing " + name, INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

// rdstd / rdstd8: land points set to 0 (input_output_data.f90:164-180).
template <class T>
int read_records(const char *path, const char *fname, int nfield, T *field, const float *lu, int nx, int ny, int nz,
                 int nxb, int nxe, int nyb, int nye, int nzb, int nze)
{
    if (!field || !lu) return fail("null pointer", INMSOM_BT_EINVAL);
    if (nfield < 1) return fail("nfield must be >= 1", INMSOM_BT_EINVAL);
    int rc = check_box(nx, ny, nz, nxb, nxe, nyb, nye, nzb, nze);
    if (rc) return rc;
    const std::string name = fulfname(path, fname);
    const int fd = ::open(name.c_str(), O_RDONLY);
    if (fd < 0) return fail("error in open file for reading: " + name + ": " + strerror(errno), INMSOM_BT_EINVAL);
    const int rx = nxe - nxb + 1, ry = nye - nyb + 1;
    const size_t recl = (size_t)rx * ry * sizeof(T);
    std::vector<T> rec((size_t)rx * ry);
    const long long nrecf = (long long)(nfield - 1) * (nze - nzb + 1);
    int kr = 0;
    for (int k = nzb; k <= nze; ++k) {
        ++kr;
        const off_t off = (off_t)(nrecf + kr - 1) * (off_t)recl;
        char *b = (char *)rec.data();
        size_t got = 0;
        while (got < recl) {
            const ssize_t r = ::pread(fd, b + got, recl - got, off + (off_t)got);
            if (r <= 0) {
                ::close(fd);
                return fail("error in reading from file: " + name + ", level " + std::to_string(k), INMSOM_BT_EINVAL);
            }
            got += (size_t)r;
        }
        size_t q = 0;
        for (int j = nyb; j <= nye; ++j)
            for (int i = nxb; i <= nxe; ++i) {
                const size_t a = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1));
                field[a] = rec[q++];
                if (std::fabs(lu[(size_t)(i - 1) + (size_t)nx * (j - 1)]) < 0.5f) field[a] = (T)0;
            }
    }
    ::close(fd);
    return INMSOM_BT_OK;
}

// ---------------------------------------------------------- Fortran editing
// Iw and Iw.m (minimum m digits); a field too narrow holds asterisks.
std::string fmt_i(long long v, int w, int m = 1)
{
    std::string d = std::to_string(v < 0 ? -v : v);
    while ((int)d.size() < m) d = "0" + d;
    if (v < 0) d = "-" + d;
    if ((int)d.size() > w) return std::string(w, '*');
    return std::string(w - d.size(), ' ') + d;
}

// Decimal digits of |x| rounded to d significant digits: "ddd..." and the
// exponent e with |x| ~ 0.ddd x 10^e (x != 0), from the exact binary value.
void sig_digits(double x, int d, std::string &digits, int &e)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", d - 1, std::fabs(x));  // d.ddde+XX
    digits.clear();
    for (const char *c = buf; *c && *c != 'e'; ++c)
        if (*c >= '0' && *c <= '9') digits.push_back(*c);
    e = std::atoi(std::strchr(buf, 'e') + 1) + 1;
}

// Ew.d: [-]0.ddddE+xx (exponent without the E when |exp| > 99).
std::string fmt_e(double x, int w, int d)
{
    std::string s;
    int e = 0;
    std::string dig;
    if (x == 0.0) {
        dig.assign(d, '0');
        e = 0;
    } else {
        sig_digits(x, d, dig, e);
    }
    char ex[16];
    if (std::abs(e) <= 99)
        std::snprintf(ex, sizeof ex, "E%c%02d", e < 0 ? '-' : '+', std::abs(e));
    else
        std::snprintf(ex, sizeof ex, "%c%03d", e < 0 ? '-' : '+', std::abs(e));
    const std::string neg = std::signbit(x) && x != 0.0 ? "-" : "";
    s = neg + "0." + dig + ex;
    if ((int)s.size() > w) s = neg + "." + dig + ex;  // the optional zero goes first
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// Fw.d
std::string fmt_f(double x, int w, int d)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "%.*f", d, x);
    std::string s = buf;
    if (d == 0) s += ".";  // Fw.0 keeps the decimal point
    if (s == "-0." + std::string(d, '0') || (s[0] == '-' && std::strspn(s.c_str() + 1, "0.") == s.size() - 1))
        s = s.substr(1);  // no minus sign on a zero result
    if ((int)s.size() > w && s.rfind("0.", 0) == 0) s = s.substr(1);
    if ((int)s.size() > w && s.rfind("-0.", 0) == 0) s = "-" + s.substr(2);
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}

// Gw.d of a real (Fortran 2008 10.7.5.2.2): F editing with 4 trailing
// blanks when 0.1 <= |x| (rounded to d digits) < 10^d, E editing otherwise.
std::string fmt_g(double x, int w, int d)
{
    if (x == 0.0) return fmt_f(0.0, w - 4, d - 1) + "    ";
    std::string dig;
    int e;
    sig_digits(x, d, dig, e);  // |x| rounded = 0.dig x 10^e
    if (e >= 0 && e <= d) return fmt_f(x, w - 4, d - e) + "    ";
    return fmt_e(x, w, d);
}

std::string pad(const std::string &s, size_t n)
{
    std::string t = s.substr(0, std::min(s.size(), n));
    t.resize(n, ' ');
    return t;
}

// sec_to_yr_mo_hr_mn (rw_ctl_file.f90:161-183) in default (4-byte) reals.
void sec_to_yr_mo_hr_mn(float sec, int &nyr, int &nmo, int &ndy, int &nhr, int &nmn)
{
    const float yr_sec = 360.0f * 86400.0f, mo_sec = 30.0f * 86400.0f;
    float sec1 = sec + 0.10f;
    nyr = (int)(sec1 / yr_sec);
    sec1 = sec1 - (float)nyr * yr_sec;
    nmo = (int)(sec1 / mo_sec);
    sec1 = sec1 - (float)nmo * mo_sec;
    ndy = (int)(sec1 / 86400.0f);
    sec1 = sec1 - (float)ndy * 86400.0f;
    nhr = (int)(sec1 / 3600.0f);
    sec1 = sec1 - (float)nhr * 3600.0f;
    nmn = (int)(sec1 / 60.0f);
}

// extname (rw_ctl_file.f90:140-158): the name's extension from its last
// '.' (not the first character) replaced by ext.
std::string extname(const std::string &name, const char *ext)
{
    size_t n = std::string::npos;
    for (size_t i = name.size(); i-- > 1;)
        if (name[i] == '.') {
            n = i;
            break;
        }
    if (n == std::string::npos) n = 1;
    return name.substr(0, n) + ext;
}

// '(a,i6,a,7(g15.8,1x)/(22x,7(g15.8,1x)))' with LEVELS, or
// '(a,i6,a,2(g15.8,5x))' with LINEAR
std::string axis_lines(const char *key, int n, int type, const double *v0, double h)
{
    std::string s = key + fmt_i(n, 6);
    if (type == 0) {
        s += "  LINEAR   " + fmt_g(v0[0], 15, 8) + "     " + fmt_g(h, 15, 8) + "\n";
        return s;
    }
    s += "  LEVELS  ";
    for (int i = 0; i < n; ++i) {
        if (i > 0 && i % 7 == 0) s += "\n" + std::string(22, ' ');
        s += fmt_g(v0[i], 15, 8);
        // 1x after a value moves the position; it writes a blank only when
        // another value follows in the same record
        if (i + 1 < n && (i + 1) % 7 != 0) s += " ";
    }
    // after the seventh value of the first record the format's '/' still
    // ends that record before the next value is looked for
    if (n == 7) s += "\n";
    s += "\n";
    return s;
}

int ctl_write(const char *fname, float undef, int nx, int ny, int nz, int nt, int xtype, const double *x0, double hx,
              int ytype, const double *y0, double hy, int ztype, const double *z0, double hz, int yr_type, int year0,
              int month0, int day0, int hour0, int minute0, float ht, const char *title, const char *varname)
{
    static const char *month[12] = {"JAN", "FEB", "MAR", "APR", "MAY", "JUN",
                                     "JUL", "AUG", "SEP", "OCT", "NOV", "DEC"};
    if (!fname || !x0 || !y0 || !z0 || !title || !varname) return fail("null pointer", INMSOM_BT_EINVAL);
    if (month0 < 1 || month0 > 12) return fail("month0 must be in 1..12", INMSOM_BT_EINVAL);
    int nyr, nmo, ndy, nhr, nmn, k = 0;
    const char *tf = nullptr;
    sec_to_yr_mo_hr_mn(ht, nyr, nmo, ndy, nhr, nmn);
    if (nyr != 0) { tf = "yr"; k = nyr; }
    else if (nmo != 0) { tf = "mo"; k = nmo; }
    else if (ndy != 0) { tf = "dy"; k = ndy; }
    else if (nhr != 0) { tf = "hr"; k = nhr; }
    else if (nmn != 0) { tf = "mn"; k = nmn; }
    if (!tf) return fail("time step error in creating ctl file (ht < 60 s)", INMSOM_BT_EINVAL);
    // namectl / namedat: character(128) copies of fname
    const std::string base = pad(fname, 128);
    std::string trimmed = base.substr(0, base.find_last_not_of(' ') + 1);
    const std::string namectl = extname(trimmed, ".ctl"), namedat = extname(trimmed, ".dat");
    const size_t slash = namedat.find_last_of('/');
    const std::string namedat2 = pad(slash == std::string::npos ? namedat : namedat.substr(slash + 1), 128);
    std::string s;
    s += "DSET    ^" + namedat2 + "\n";
    s += std::string("TITLE    ") + title + "\n";
    s += "UNDEF   " + fmt_e(undef, 12, 5) + "  ! gap value\n";
    s += axis_lines("XDEF  ", nx, xtype, x0, hx);
    s += axis_lines("YDEF  ", ny, ytype, y0, hy);
    s += axis_lines("ZDEF  ", nz, ztype, z0, hz);
    s += "TDEF  " + fmt_i(nt, 6) + "  LINEAR  " + fmt_i(hour0, 2, 2) + ":" + fmt_i(minute0, 2, 2) + "Z" +
         fmt_i(day0, 2, 2) + month[month0 - 1] + fmt_i(year0, 4, 4) + "     " + fmt_i(k, 4) + tf + "\n";
    if (yr_type == 0) s += "OPTIONS  365_DAY_CALENDAR\n";
    s += "VARS 1  ! Number of variables\n";
    s += std::string(varname) + "        " + fmt_i(nz, 5) + "  1 VARIABLES \n";
    s += "ENDVARS\n";
    s += "\n";
    FILE *f = std::fopen(namectl.c_str(), "wb");
    if (!f) return fail("cannot open " + namectl + ": " + strerror(errno), INMSOM_BT_EINVAL);
    const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
    if (std::fclose(f) != 0 || !ok) return fail("error writing " + namectl, INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

}  // namespace

int write_record(const char *path, const char *fname, long long rec, const void *data, size_t recl)
{
    const std::string name = fulfname(path, fname);
    const int fd = ::open(name.c_str(), O_WRONLY | O_CREAT, 0644);
    if (fd < 0) return fail("error in open file for writing: " + name + ": " + strerror(errno), INMSOM_BT_EINVAL);
    const char *b = (const char *)data;
    const off_t off = (off_t)(rec - 1) * (off_t)recl;
    size_t left = recl;
    while (left > 0) {
        const ssize_t w = ::pwrite(fd, b, left, off + (off_t)(recl - left));
        if (w <= 0) {
            ::close(fd);
            return fail("error in writing on file: " + name, INMSOM_BT_EINVAL);
        }
        b += w;
        left -= (size_t)w;
    }
    if (::close(fd) != 0) return fail("error closing " + name, INMSOM_BT_EINVAL);
    return INMSOM_BT_OK;
}

int write_ctl(const char *path, const char *fname, float undef, int nx, int ny, int nt, double x0, double hx,
              double y0, double hy, int yr_type, int year0, int month0, int day0, int hour0, int minute0, float ht,
              const char *title, const char *varname)
{
    const std::string name = fulfname(path, fname);
    const double z0 = 0.0;
    return ctl_write(name.c_str(), undef, nx, ny, 1, nt, 0, &x0, hx, 0, &y0, hy, 0, &z0, 1.0, yr_type, year0, month0,
                     day0, hour0, minute0, ht, title, varname);
}

}  // namespace io
}  // namespace ibt

using namespace ibt::io;

extern "C" {

int inmsom_bt_wdstd(const char *path, const char *fname, int nfield, float *field, const float *lu, int nx, int ny,
                    int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze)
{
    return write_records<float>(path, fname, nfield, field, lu, nx, ny, nz, nxb, nxe, nyb, nye, nzb, nze);
}

int inmsom_bt_rdstd(const char *path, const char *fname, int nfield, float *field, const float *lu, int nx, int ny,
                    int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze)
{
    return read_records<float>(path, fname, nfield, field, lu, nx, ny, nz, nxb, nxe, nyb, nye, nzb, nze);
}

int inmsom_bt_wdstd8(const char *path, const char *fname, int nfield, double *field, const float *lu, int nx, int ny,
                     int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze)
{
    return write_records<double>(path, fname, nfield, field, lu, nx, ny, nz, nxb, nxe, nyb, nye, nzb, nze);
}

int inmsom_bt_rdstd8(const char *path, const char *fname, int nfield, double *field, const float *lu, int nx, int ny,
                     int nz, int nxb, int nxe, int nyb, int nye, int nzb, int nze)
{
    return read_records<double>(path, fname, nfield, field, lu, nx, ny, nz, nxb, nxe, nyb, nye, nzb, nze);
}

int inmsom_bt_ctl_write(const char *fname, float undef, int nx, int ny, int nz, int nt, int xtype, const double *x0,
                        double hx, int ytype, const double *y0, double hy, int ztype, const double *z0, double hz,
                        int yr_type, int year0, int month0, int day0, int hour0, int minute0, float ht,
                        const char *title, const char *varname)
{
    return ctl_write(fname, undef, nx, ny, nz, nt, xtype, x0, hx, ytype, y0, hy, ztype, z0, hz, yr_type, year0,
                     month0, day0, hour0, minute0, ht, title, varname);
}

}  // extern "C"
