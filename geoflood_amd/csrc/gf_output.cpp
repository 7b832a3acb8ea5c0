// gf_output.cpp -- ASCII frame and gauge record formatting (host code).
//
// The byte layout is the reference's, because post-processing tools read
// these files by position: fort.tNNNN / fort.qNNNN as written by
// fc2d_geoclaw_output_ascii_fort.f:1-117 with Flang's formatted output, and
// gaugeNNNNN.txt as written by fc2d_geoclaw_gauges_default.c:172-200 and
// :275-305.
#include "gf_output.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace gf {
namespace ascii {

void format_E(char* out, double v, int w, int d)
{
    char body[64];
    int n = 0;
    if (std::isnan(v)) {
        n = std::snprintf(body, sizeof body, "NaN");
    } else if (std::isinf(v)) {
        n = std::snprintf(body, sizeof body, v < 0 ? "-Inf" : "Inf");
    } else {
        // C's %e rounds to nearest with ties to even on the exact binary
        // value, as Flang's E editing does; only the layout differs
        // (d.ddd e+x  ->  0.dddd E+(x+1)).
        char mant[40];
        int e10 = 0;
        if (v == 0.0) {
            std::memset(mant, '0', (size_t)d);
        } else {
            char t[64];
            std::snprintf(t, sizeof t, "%.*e", d - 1, std::fabs(v));
            mant[0] = t[0];
            std::memcpy(mant + 1, t + 2, (size_t)(d - 1));
            e10 = std::atoi(std::strchr(t, 'e') + 1) + 1;
        }
        mant[d] = 0;
        const int ae = e10 < 0 ? -e10 : e10;
        const char sign = e10 < 0 ? '-' : '+';
        const char* neg = std::signbit(v) ? "-" : "";
        n = ae <= 99 ? std::snprintf(body, sizeof body, "%s0.%sE%c%02d", neg, mant, sign, ae)
                     : std::snprintf(body, sizeof body, "%s0.%s%c%03d", neg, mant, sign, ae);
    }
    if (n > w) {
        std::memset(out, '*', (size_t)w);
        return;
    }
    std::memset(out, ' ', (size_t)(w - n));
    std::memcpy(out + (w - n), body, (size_t)n);
}

static void put_E(std::string& s, double v, int w, int d)
{
    char t[64];
    format_E(t, v, w, d);
    s.append(t, (size_t)w);
}

// Fortran I5 followed by the label of the reference's format statements.
static void put_i5_line(std::string& s, int v, const char* label)
{
    char t[32];
    const int n = std::snprintf(t, sizeof t, "%d", v);
    if (n > 5)
        s.append("*****");
    else
        s.append((size_t)(5 - n), ' ').append(t, (size_t)n);
    s.append(17, ' ').append(label).push_back('\n');
}

void frame_header(std::string& s, double time, int meqn, int maux, int ngrids, int mbc)
{
    put_E(s, time, 30, 20);
    s.append("    time\n");
    put_i5_line(s, meqn + 1, "meqn");
    put_i5_line(s, ngrids, "ngrids");
    put_i5_line(s, maux, "num_aux");
    put_i5_line(s, 2, "num_dim");
    put_i5_line(s, mbc, "num_ghost");
}

void frame_patch(std::string& s, const double* v, int mx, int my, double xlower, double ylower, double dx,
                 double dy, int patch_num, int level, int blockno, int mpirank)
{
    put_i5_line(s, patch_num, "grid_number");
    put_i5_line(s, level, "AMR_level");
    put_i5_line(s, blockno, "block_number");
    put_i5_line(s, mpirank, "mpi_rank");
    put_i5_line(s, mx, "mx");
    put_i5_line(s, my, "my");
    put_E(s, xlower, 24, 16);
    s.append("    xlow\n");
    put_E(s, ylower, 24, 16);
    s.append("    ylow\n");
    put_E(s, dx, 24, 16);
    s.append("    dx\n");
    put_E(s, dy, 24, 16);
    s.append("    dy\n\n");
    char line[4 * 26 + 1];
    for (int j = 0; j < my; j++) {
        for (int i = 0; i < mx; i++) {
            const double* c = v + 4 * ((size_t)j * mx + i);
            for (int m = 0; m < 4; m++) format_E(line + 26 * m, c[m], 26, 16);
            line[4 * 26] = '\n';
            s.append(line, sizeof line);
        }
        s.append("  \n");  // list-directed write of ' ': a leading blank and the blank
    }
}

void gauge_header(std::string& s, int id, double xc, double yc)
{
    char t[256];
    const int n = std::snprintf(t, sizeof t,
                                "# gauge_id= %5d location=( %17.10e %17.10e ) num_eqn= %2d\n"
                                "# Columns: level time h    hu    hv    eta\n",
                                id, xc, yc, 4);
    s.append(t, (size_t)n);
}

void gauge_line(std::string& s, int level, double t, double h, double hu, double hv, double b)
{
    double eta = h + b;
    eta = std::fabs(eta) < 1e-99 ? 0 : eta;
    char l[256];
    const int n = std::snprintf(l, sizeof l, "%5d %15.7e %15.7e %15.7e %15.7e %15.7e\n", level, t, h, hu, hv, eta);
    s.append(l, (size_t)n);
}

}  // namespace ascii
}  // namespace gf
