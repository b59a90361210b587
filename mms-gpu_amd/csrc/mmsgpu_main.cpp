// mmsgpu_main.cpp -- the simulator's command-line driver over libmmsgpu.so.
//
// Same 18 positional arguments, defaults, parameter printout, geometry,
// initial conditions and output files as the reference's main (main.cpp:21-268):
//
//   mmsgpu imagefile P finaltime timestep sample_width sample_height
//          sample_depth xdim ydim zdim demag exchange external use_fmm
//          use_gpu sim_name seed IC
//
//   - imagefile: an 8-bit PNG mask (load_mask, helper_functions.cpp:173-219:
//     grey = (byte)(.3 r + .59 g + .11 b), scanline 0 is the bottom image
//     row, a cell is magnetic where grey == 0); its size replaces xdim, ydim.
//     A name that is not an existing file selects the built-in mask of the
//     reference's build without an image library (main.cpp:120-129: every
//     cell but the 1-cell border), which its logged runs used.
//   - layers 1 .. zdim-2 of the magnetic mask carry material (main.cpp:192-193);
//   - IC 0 single domain along x, 1 four-quadrant vortex, 2 random in-plane
//     angles from srand(seed) / rand() (main.cpp:144-227, numerics.cu:125-130);
//   - writes <sim_name>/M.dat (save_vector3d) and <sim_name>/material.dat
//     (layer 1 as a matrix, matrix2file), then runs time_marching, which
//     writes dynamics.dat, Mdynamics.dat and panic.dat.
//
// Differences: P may be 1..9 (the reference asserts P <= 4); grey and
// grey+alpha PNGs use their grey value; the pause for a key press after the
// first step is not reproduced.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <time.h>
#include <zlib.h>

#include <string>
#include <vector>

#include "../../include/mms_gpu.h"

namespace {

struct Vec3 {
    float x, y, z;
};

unsigned be32(const unsigned char *p) { return (unsigned)p[0] << 24 | (unsigned)p[1] << 16 | (unsigned)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Decodes a non-interlaced 8-bit PNG (grey, grey + alpha, RGB, RGBA,
// palette) into grey levels, bottom image row first.
bool load_mask_png(const char *path, std::vector<unsigned char> &grey, int *w_out, int *h_out) {
    FILE *fh = fopen(path, "rb");
    if (!fh) {
        fprintf(stderr, "ERROR: cannot open mask file %s\n", path);
        return false;
    }
    std::vector<unsigned char> data;
    unsigned char buf[65536];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, fh)) > 0) data.insert(data.end(), buf, buf + got);
    fclose(fh);
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (data.size() < 8 || memcmp(data.data(), sig, 8) != 0) {
        fprintf(stderr, "ERROR: %s is not a PNG file\n", path);
        return false;
    }
    unsigned w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat, palette;
    size_t pos = 8;
    while (pos + 8 <= data.size()) {
        const unsigned len = be32(&data[pos]);
        if (pos + 12 + (size_t)len > data.size()) break;
        const unsigned char *type = &data[pos + 4], *body = &data[pos + 8];
        if (!memcmp(type, "IHDR", 4) && len >= 13) {
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (!memcmp(type, "PLTE", 4)) {
            palette.assign(body, body + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    const int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (depth != 8 || interlace != 0 || ch == 0 || w == 0 || h == 0 || (ctype == 3 && palette.empty())) {
        fprintf(stderr, "ERROR: %s: only 8-bit non-interlaced PNGs are supported\n", path);
        return false;
    }
    const size_t stride = (size_t)w * ch;
    std::vector<unsigned char> raw(h * (stride + 1));
    uLongf rawlen = raw.size();
    if (uncompress(raw.data(), &rawlen, idat.data(), idat.size()) != Z_OK || rawlen != raw.size()) {
        fprintf(stderr, "ERROR: %s: corrupt image data\n", path);
        return false;
    }
    std::vector<unsigned char> img(h * stride), prev(stride, 0);
    for (unsigned r = 0; r < h; r++) {
        const unsigned char f = raw[r * (stride + 1)], *line = &raw[r * (stride + 1) + 1];
        unsigned char *cur = &img[r * stride];
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= (size_t)ch ? cur[i - ch] : 0, b = prev[i], c = i >= (size_t)ch ? prev[i - ch] : 0;
            int v;
            switch (f) {
            case 0: v = line[i]; break;
            case 1: v = line[i] + a; break;
            case 2: v = line[i] + b; break;
            case 3: v = line[i] + (a + b) / 2; break;
            case 4: v = line[i] + paeth(a, b, c); break;
            default: fprintf(stderr, "ERROR: %s: bad row filter %d\n", path, f); return false;
            }
            cur[i] = (unsigned char)(v & 0xFF);
        }
        memcpy(prev.data(), cur, stride);
    }
    grey.assign((size_t)w * h, 0);
    for (unsigned r = 0; r < h; r++)
        for (unsigned x = 0; x < w; x++) {
            const unsigned char *px = &img[r * stride + (size_t)x * ch];
            unsigned char g;
            if (ctype == 0 || ctype == 4) {
                g = px[0];
            } else {
                const unsigned char *rgb = (ctype == 3) ? &palette[3 * (size_t)px[0]] : px;
                const double rr = rgb[0], gg = rgb[1], bb = rgb[2];
                g = (unsigned char)(.3 * rr + .59 * gg + .11 * bb);
            }
            grey[(size_t)(h - 1 - r) * w + x] = g;  // scanline 0 = bottom row
        }
    *w_out = (int)w;
    *h_out = (int)h;
    return true;
}

// numerics.cu:125-130
float frand_atob(const float a, const float b) {
    double r = rand() / (double)RAND_MAX;
    r = a + (b - a) * r;
    return (float)r;
}

// helper_functions.cpp:80-103 (x outer, y inner within each layer)
int save_vector3d(const std::vector<Vec3> &v, int zdim, int ydim, int xdim, const char *filename) {
    FILE *fh = fopen(filename, "w");
    if (!fh) {
        fprintf(stderr, "FATAL ERROR: couldn't open file %s\n", filename);
        return EXIT_FAILURE;
    }
    for (int z = 0; z < zdim; z++)
        for (int x = 0; x < xdim; x++)
            for (int y = 0; y < ydim; y++) {
                const Vec3 &a = v[(size_t)z * ydim * xdim + (size_t)y * xdim + x];
                fprintf(fh, "%g %g %g \n", a.x, a.y, a.z);
            }
    fclose(fh);
    return EXIT_SUCCESS;
}

// helper_functions.cpp:33-51
int matrix2file(const std::vector<float> &m, int rows, int cols, const char *filename) {
    FILE *fh = fopen(filename, "w");
    if (!fh) {
        printf("FATAL ERROR: Error opening file %s\n", filename);
        return EXIT_FAILURE;
    }
    for (int r = 0; r < rows; r++) {
        for (int c = 0; c < cols; c++) fprintf(fh, "%g ", m[(size_t)r * cols + c]);
        fprintf(fh, "\n");
    }
    fclose(fh);
    return EXIT_SUCCESS;
}

bool ends_with_png(const char *s) {
    const size_t n = strlen(s);
    return n >= 3 && (!strcmp(s + n - 3, "png") || !strcmp(s + n - 3, "PNG"));
}

}  // namespace

int main(int argc, char **argv) {
    int status = 0;
    const int verbosity = 4;
    timeval time1, time2;
    status |= gettimeofday(&time1, NULL);

    // defaults (main.cpp:30-47)
    char filename_arg[1000] = "verysmall_16x16.png";
    unsigned int P = 3;
    float finaltime = 1e-9f, timestep = 1e-14f;
    float sample_width = 10e-9f, sample_height = 10e-9f, sample_depth = 1e-9f;
    int xdim = 16, ydim = 16, zdim = 3;
    int demag = 1, exchange = 1, external = 0, use_fmm = 0, use_gpu = 0;
    char sim_name[1000] = "sim_untitled";
    unsigned int seed = (unsigned int)time(NULL);
    unsigned int IC = 0;  // 0 single domain, 1 vortex, 2 random
    // positional arguments (main.cpp:49-70)
    if (argc >= 2) sscanf(argv[1], "%999s", filename_arg);
    if (argc >= 3) {
        sscanf(argv[2], "%u", &P);
        if (P > 9) {
            fprintf(stderr, "ERROR: P must be at most 9\n");
            return EXIT_FAILURE;
        }
    }
    if (argc >= 4) sscanf(argv[3], "%f", &finaltime);
    if (argc >= 5) sscanf(argv[4], "%f", &timestep);
    if (argc >= 6) sscanf(argv[5], "%f", &sample_width);
    if (argc >= 7) sscanf(argv[6], "%f", &sample_height);
    if (argc >= 8) sscanf(argv[7], "%f", &sample_depth);
    if (argc >= 9) sscanf(argv[8], "%d", &xdim);
    if (argc >= 10) sscanf(argv[9], "%d", &ydim);
    if (argc >= 11) sscanf(argv[10], "%d", &zdim);
    if (argc >= 12) sscanf(argv[11], "%d", &demag);
    if (argc >= 13) sscanf(argv[12], "%d", &exchange);
    if (argc >= 14) sscanf(argv[13], "%d", &external);
    if (argc >= 15) sscanf(argv[14], "%d", &use_fmm);
    if (argc >= 16) sscanf(argv[15], "%d", &use_gpu);
    if (argc >= 17) sscanf(argv[16], "%999s", sim_name);
    if (argc >= 18) sscanf(argv[17], "%u", &seed);
    if (argc >= 19) sscanf(argv[18], "%u", &IC);
    srand(seed);
    if (verbosity >= 2) {
        printf("imagefile = %s \n", filename_arg);
        printf("P = %d \n", P);
        printf("finaltime = %g \n", finaltime);
        printf("timestep = %g \n", timestep);
        printf("sample_width  = %g \n", sample_width);
        printf("sample_height = %g \n", sample_height);
        printf("sample_depth  = %g \n", sample_depth);
        printf("xdim = %d \n", xdim);
        printf("ydim = %d \n", ydim);
        printf("zdim = %d \n", zdim);
        printf("demag = %d \n", demag);
        printf("exchange = %d \n", exchange);
        printf("external = %d \n", external);
        printf("use_fmm = %d \n", use_fmm);
        printf("use_gpu = %d \n", use_gpu);
        printf("sim_name = %s \n", sim_name);
        printf("SEED = %d \n", seed);
        printf("IC = %d \n", IC);
    }
    mkdir(sim_name, 0755);
    char filename[1100];

    // material constants (main.cpp:107-111)
    const float mu_0 = 4 * M_PI * 1e-7, Ms = 8.6e5, Aexch = 1.3e-11, alfa = 0.5, gamma = 2.21e5;

    // mask: grey level per in-plane cell, material where 0
    std::vector<unsigned char> mask;
    struct stat sb;
    if (stat(filename_arg, &sb) == 0) {
        if (!ends_with_png(filename_arg)) {
            fprintf(stderr, "ERROR: the mask file %s is not a .png file\n", filename_arg);
            return EXIT_FAILURE;
        }
        if (!load_mask_png(filename_arg, mask, &xdim, &ydim)) return EXIT_FAILURE;
    } else {
        printf("mask file %s not found: built-in mask (all cells but the 1-cell border)\n", filename_arg);
        if (xdim <= 0 || ydim <= 0) {
            fprintf(stderr, "ERROR: xdim and ydim must be positive\n");
            return EXIT_FAILURE;
        }
        mask.assign((size_t)xdim * ydim, 1);
        for (int y = 1; y < ydim - 1; y++)
            for (int x = 1; x < xdim - 1; x++) mask[(size_t)y * xdim + x] = 0;
    }
    const float dx = sample_width / xdim, dy = sample_height / ydim, dz = sample_depth / zdim;
    printf("(xdim, ydim, zdim) = (%d, %d, %d)\n", xdim, ydim, zdim);
    printf("(sample_width, sample_height, sample_depth) = (%g, %g, %g)\n", sample_width, sample_height, sample_depth);
    printf("(dx, dy, dz) = (%g, %g, %g)\n", dx, dy, dz);

    switch (IC) {
    case 0: printf("Initial condition: Single Domain\n"); break;
    case 1: printf("Initial condition: Vortex\n"); break;
    case 2: printf("Initial condition: Random\n"); break;
    default: fprintf(stderr, "ERROR: Unknown Initial Condition!\n"); return EXIT_FAILURE;
    }
    if (zdim < 3) {
        fprintf(stderr, "ERROR: zdim must be at least 3 (layers 1..zdim-2 are magnetic)\n");
        return EXIT_FAILURE;
    }

    // initial magnetisation (main.cpp:167-227)
    const size_t n = (size_t)zdim * ydim * xdim;
    std::vector<unsigned char> material(n, 0);
    std::vector<Vec3> M(n, Vec3{0.0f, 0.0f, 0.0f});
    for (int z = 1; z < zdim - 1; z++)
        for (int y = 0; y < ydim; y++)
            for (int x = 0; x < xdim; x++) {
                if (mask[(size_t)y * xdim + x]) continue;
                const float theta = M_PI / 2;
                float phi = 0;
                if (IC == 0) {
                    phi = 0 * M_PI / 180;
                } else if (IC == 2) {
                    phi = frand_atob(0, 360) * M_PI / 180;
                } else {
                    if ((x - xdim / 2 > 0) && (y - ydim / 2 > 0)) phi = -M_PI / 4;
                    else if ((x - xdim / 2 <= 0) && (y - ydim / 2 > 0)) phi = M_PI / 4;
                    else if ((x - xdim / 2 <= 0) && (y - ydim / 2 <= 0)) phi = 3 * M_PI / 4;
                    else if ((x - xdim / 2 > 0) && (y - ydim / 2 <= 0)) phi = -3 * M_PI / 4;
                }
                // the unit vector in double (sin/cos of the float angles),
                // each component rounded to float, then scaled by Ms in float
                const float ux = (float)(sin((double)theta) * cos((double)phi));
                const float uy = (float)(sin((double)theta) * sin((double)phi));
                const float uz = (float)cos((double)theta);
                const size_t i = (size_t)z * ydim * xdim + (size_t)y * xdim + x;
                M[i] = Vec3{Ms * ux, Ms * uy, Ms * uz};
                material[i] = 1;
            }

    snprintf(filename, sizeof filename, "%s/%s", sim_name, "M.dat");
    status |= save_vector3d(M, zdim, ydim, xdim, filename);
    if (status) return EXIT_FAILURE;
    {
        std::vector<float> m((size_t)xdim * ydim);
        const int z = 1;
        for (int y = 0; y < ydim; y++)
            for (int x = 0; x < xdim; x++)
                m[(size_t)y * xdim + x] = (float)material[(size_t)z * ydim * xdim + (size_t)y * xdim + x];
        snprintf(filename, sizeof filename, "%s/%s", sim_name, "material.dat");
        status |= matrix2file(m, ydim, xdim, filename);
        if (status) return EXIT_FAILURE;
    }

    if (time_marching(material.data(), &M[0].x, finaltime, timestep, xdim, ydim, zdim, dx, dy, dz, (int)P, mu_0, Ms,
                      Aexch, alfa, gamma, demag, exchange, external, use_fmm, use_gpu, sim_name, verbosity)) {
        fprintf(stderr, "ERROR: time_marching: %s\n", mms_last_error());
        status = 1;
    }
    printf("SEED = %d\n", seed);
    status |= gettimeofday(&time2, NULL);
    const double deltatime = (time2.tv_sec + time2.tv_usec / 1e6) - (time1.tv_sec + time1.tv_usec / 1e6);
    printf("Simulation completed in %f seconds.\n", deltatime);
    return status ? EXIT_FAILURE : EXIT_SUCCESS;
}
