// gf_output.h -- host-side ASCII formatting of frames and gauge records in
// the reference's file formats (fc2d_geoclaw_output_ascii_fort.f,
// fc2d_geoclaw_gauges_default.c).
#pragma once

#include <string>

namespace gf {
namespace ascii {

// Fortran Ew.d (Flang): [-]0.<d digits>E+xx (or +xxx for a three-digit
// exponent), Inf / -Inf / NaN, right-justified in exactly w characters.
void format_E(char* out, double v, int w, int d);

// fort.tNNNN (fc2d_geoclaw_fort_write_header, format 1000).
void frame_header(std::string& s, double time, int meqn, int maux, int ngrids, int mbc);

// One patch's block of fort.qNNNN (fc2d_geoclaw_fort_write_file, formats
// 1001, 1002, 120): v holds mx*my cells in row order, four values each
// (h, hu, hv, eta), already flushed below 1e-99.
void frame_patch(std::string& s, const double* v, int mx, int my, double xlower, double ylower, double dx,
                 double dy, int patch_num, int level, int blockno, int mpirank);

// Gauge file header lines and one record (geoclaw_create_gauge_files_default,
// geoclaw_print_gauges_default).
void gauge_header(std::string& s, int id, double xc, double yc);
void gauge_line(std::string& s, int level, double t, double h, double hu, double hv, double b);

}  // namespace ascii
}  // namespace gf
