// Internal interface of the file-format writers (bt_io.cpp) for the
// device snapshot (bt_snapshot.hip).
#pragma once

#include <cstddef>

namespace ibt {
namespace io {

// record `rec` (1-based) of recl bytes of the DIOGIN file path/fname
int write_record(const char *path, const char *fname, long long rec, const void *data, size_t recl);
// GrADS descriptor of path/fname with linear x, y axes and one level
int write_ctl(const char *path, const char *fname, float undef, int nx, int ny, int nt, double x0, double hx,
              double y0, double hy, int yr_type, int year0, int month0, int day0, int hour0, int minute0, float ht,
              const char *title, const char *varname);

}  // namespace io
}  // namespace ibt
