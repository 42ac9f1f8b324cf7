// ref_harness.cpp -- a thin C entry layer over the reference's own kriging
// core, compiled together with the reference sources (never copied) into
// oracle/_ref/libhpgl_ref.so by oracle/Makefile.  TEST INFRASTRUCTURE ONLY:
// it is loaded by tests/ (and the CPU-baseline leg of bench.py) as a checker
// and as the measured reference CPU path; the product never links it.
//
// It calls the reference drivers directly (bypassing api.cpp, which drags in
// the simulation and file-I/O code) and also exposes the reference's own
// neighbour lookup, so neighbour sets can be compared element by element.
#include "stdafx.h"

#include "api.h"
#include "property_array.h"
#include "sugarbox_grid.h"
#include "ok_params.h"
#include "sk_params.h"
#include "output.h"
#include "pretty_printer.h"
#include "cont_kriging.h"
#include "cov_model.h"
#include "covariance_field.h"
#include "hpgl_core.h"
#include "ik_params.h"
#include "median_ik.h"
#include "my_kriging_weights.h"
#include "ok_params.h"
#include "precalculated_covariance.h"
#include "property_writer.h"
#include "property_array.h"
#include "sk_params.h"
#include "sugarbox_grid.h"
#include "sugarbox_indexed_neighbour_lookup.h"

#include <omp.h>

namespace {

int quiet_handler(char*, void*) { return 0; }

template <typename P>
void fill_cov(P& p, int type, const double* ranges, const double* angles, double sill, double nugget) {
    p.m_covariance_type = (hpgl::covariance_type_t)type;
    p.set_ranges(ranges[0], ranges[1], ranges[2]);
    p.set_angles(angles[0], angles[1], angles[2]);
    p.m_sill = sill;
    p.m_nugget = nugget;
}

template <typename cov_t>
void dump_neighbours(const hpgl::sugarbox_grid_t& grid, const cov_t& cov, const hpgl::neighbourhood_param_t& nbp,
                     const hpgl::cont_property_array_t& prop, int all_cells, int* counts, int* idx_out, int kmax) {
    typedef hpgl::indexed_neighbour_lookup_t<hpgl::sugarbox_grid_t, cov_t> nl_t;
    nl_t nl(&grid, &cov, nbp);
    for (int node = 0; node < prop.size(); ++node)
        if (prop.is_informed(node)) nl.add_node(node);
    hpgl::is_informed_predicate_t<hpgl::cont_property_array_t> pred(prop);
#pragma omp parallel for schedule(dynamic, 256)
    for (int node = 0; node < prop.size(); ++node) {
        std::vector<hpgl::node_index_t> indices;
        std::vector<hpgl::sugarbox_location_t> coords;
        hpgl::sugarbox_location_t c;
        for (int i = 0; i < kmax; ++i) idx_out[(size_t)node * kmax + i] = -1;
        counts[node] = 0;
        if (!all_cells && prop.is_informed(node)) continue;
        nl.find(node, pred, c, indices, coords);
        counts[node] = (int)indices.size();
        for (size_t i = 0; i < indices.size() && (int)i < kmax; ++i) idx_out[(size_t)node * kmax + i] = indices[i];
    }
}

}  // namespace

extern "C" {

void ref_set_threads(int n) {
    omp_set_num_threads(n);
}

void ref_quiet() {
    hpgl_set_output_handler(quiet_handler, nullptr);
}

// hpgl::ordinary_kriging on float data / byte mask, output buffers pre-filled
// by the caller with a clone of the input (as geo.py does).
int ref_ordinary_kriging(float* data, unsigned char* mask, int nx, int ny, int nz, int type, const double* ranges,
                         const double* angles, double sill, double nugget, const int* radiuses, int max_nb,
                         float* out_data, unsigned char* out_mask) {
    try {
        int n = nx * ny * nz;
        hpgl::cont_property_array_t in(data, mask, n), out(out_data, out_mask, n);
        hpgl::sugarbox_grid_t grid;
        grid.init(nx, ny, nz);
        hpgl::ok_params_t p;
        fill_cov(p, type, ranges, angles, sill, nugget);
        p.set_radiuses(radiuses[0], radiuses[1], radiuses[2]);
        p.m_max_neighbours = max_nb;
        hpgl::ordinary_kriging(in, grid, p, out, true);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_simple_kriging(float* data, unsigned char* mask, int nx, int ny, int nz, int type, const double* ranges,
                       const double* angles, double sill, double nugget, const int* radiuses, int max_nb,
                       int auto_mean, double mean, float* out_data, unsigned char* out_mask) {
    try {
        int n = nx * ny * nz;
        hpgl::cont_property_array_t in(data, mask, n), out(out_data, out_mask, n);
        hpgl::sugarbox_grid_t grid;
        grid.init(nx, ny, nz);
        hpgl::sk_params_t p;
        fill_cov(p, type, ranges, angles, sill, nugget);
        p.set_radiuses(radiuses[0], radiuses[1], radiuses[2]);
        p.m_max_neighbours = max_nb;
        if (!auto_mean) p.set_mean(mean);
        hpgl::simple_kriging(in, grid, p, out);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_median_ik(unsigned char* data, unsigned char* mask, int nx, int ny, int nz, int type, const double* ranges,
                  const double* angles, double sill, double nugget, const int* radiuses, int max_nb, double p0,
                  double p1, unsigned char* out_data, unsigned char* out_mask) {
    try {
        int n = nx * ny * nz;
        hpgl::indicator_property_array_t in(data, mask, n, 2), out(out_data, out_mask, n, 2);
        hpgl::sugarbox_grid_t grid;
        grid.init(nx, ny, nz);
        hpgl::median_ik_params p;
        fill_cov(p, type, ranges, angles, sill, nugget);
        p.set_radiuses(radiuses[0], radiuses[1], radiuses[2]);
        p.m_max_neighbours = max_nb;
        p.m_marginal_probs[0] = p0;
        p.m_marginal_probs[1] = p1;
        hpgl::median_ik_for_two_indicators(p, grid, in, out);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// hpgl::lvm_kriging (the path behind hpgl_lvm_kriging): mean_data holds one
// mean per cell; output buffers pre-filled with a clone of the input.
int ref_lvm_kriging(float* data, unsigned char* mask, float* mean_data, int nx, int ny, int nz, int type,
                    const double* ranges, const double* angles, double sill, double nugget, const int* radiuses,
                    int max_nb, float* out_data, unsigned char* out_mask) {
    try {
        int n = nx * ny * nz;
        hpgl::cont_property_array_t in(data, mask, n), out(out_data, out_mask, n);
        hpgl::sugarbox_grid_t grid;
        grid.init(nx, ny, nz);
        hpgl::ok_params_t p;
        fill_cov(p, type, ranges, angles, sill, nugget);
        p.set_radiuses(radiuses[0], radiuses[1], radiuses[2]);
        p.m_max_neighbours = max_nb;
        hpgl::lvm_kriging(in, mean_data, grid, p, out);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// hpgl::indicator_kriging (the path behind hpgl_indicator_kriging) with
// `ncat` categories; per category c: types[c], ranges/angles[3c..3c+2],
// sills[c], nuggets[c], radiuses[3c..3c+2], max_nb[c], marg[c].
int ref_indicator_kriging(unsigned char* data, unsigned char* mask, int nx, int ny, int nz, int ncat,
                          const int* types, const double* ranges, const double* angles, const double* sills,
                          const double* nuggets, const int* radiuses, const int* max_nb, const double* marg,
                          unsigned char* out_data, unsigned char* out_mask) {
    try {
        int n = nx * ny * nz;
        hpgl::indicator_property_array_t in(data, mask, n, ncat), out(out_data, out_mask, n, ncat);
        hpgl::sugarbox_grid_t grid;
        grid.init(nx, ny, nz);
        hpgl::ik_params_t p;
        for (int c = 0; c < ncat; ++c)
            p.add_indicator((hpgl::covariance_type_t)types[c], ranges[3 * c], ranges[3 * c + 1], ranges[3 * c + 2],
                            angles[3 * c], angles[3 * c + 1], angles[3 * c + 2], sills[c], nuggets[c],
                            radiuses[3 * c], radiuses[3 * c + 1], radiuses[3 * c + 2], (size_t)max_nb[c], marg[c]);
        hpgl::indicator_kriging(in, grid, p, out);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// hpgl::read_inc_file_float / read_inc_file_byte (read_inc_file.cpp:137-193),
// the readers behind hpgl_read_inc_file_float / _byte (the byte form
// without api.cpp's value remapping).  Returns 0, or -1 when they throw.
int ref_read_inc_float(const char* filename, float undef, int size, float* data, unsigned char* mask) {
    try {
        hpgl::read_inc_file_float(filename, undef, size, data, mask);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_read_inc_byte(const char* filename, int undef, int size, unsigned char* data, unsigned char* mask) {
    try {
        hpgl::read_inc_file_byte(filename, undef, size, data, mask);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// property_writer_t (property_writer.cpp:44-139): INC (gslib = 0) or GSLIB
// (gslib = 1) text of a float property, or of a byte property with `remap`.
int ref_write_float(const char* filename, const char* name, float* data, unsigned char* mask, int n, double undef,
                    int gslib) {
    try {
        hpgl::property_writer_t w;
        w.init(filename, name);
        if (gslib)
            w.write_gslib_double(std::make_shared<hpgl::cont_property_array_t>(data, mask, n), undef);
        else
            w.write_double(hpgl::cont_property_array_t(data, mask, n), undef);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_write_byte(const char* filename, const char* name, unsigned char* data, unsigned char* mask, int n,
                   int indicator_count, unsigned char undef, const unsigned char* remap, int remap_count, int gslib) {
    try {
        hpgl::property_writer_t w;
        w.init(filename, name);
        std::vector<unsigned char> table(remap, remap + remap_count);
        if (gslib)
            w.write_gslib_byte(std::make_shared<hpgl::indicator_property_array_t>(data, mask, n, indicator_count),
                               undef, table);
        else
            w.write_byte(hpgl::indicator_property_array_t(data, mask, n, indicator_count), undef, table);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Neighbour lists from the reference's own lookup classes, built the way each
// driver builds them.  kind: 0 OK (cov_model_t), 1 SK (precalculated table),
// 2 MIK (covariance field, all cells searched).
int ref_neighbours(int kind, unsigned char* mask, int nx, int ny, int nz, int type, const double* ranges,
                   const double* angles, double sill, double nugget, const int* radiuses, int max_nb, int* counts,
                   int* idx_out) {
    try {
        int n = nx * ny * nz;
        std::vector<float> dummy(n, 0.0f);
        hpgl::cont_property_array_t prop(dummy.data(), mask, n);
        hpgl::sugarbox_grid_t grid;
        grid.init(nx, ny, nz);
        hpgl::ok_params_t p;
        fill_cov(p, type, ranges, angles, sill, nugget);
        p.set_radiuses(radiuses[0], radiuses[1], radiuses[2]);
        p.m_max_neighbours = max_nb;
        hpgl::cov_model_t model(p);
        if (kind == 0) {
            dump_neighbours(grid, model, p, prop, 0, counts, idx_out, max_nb);
        } else if (kind == 1) {
            hpgl::precalculated_covariances_t pcov(model, p.m_radiuses);
            dump_neighbours(grid, pcov, p, prop, 0, counts, idx_out, max_nb);
        } else {
            hpgl::covariance_field_t field(p.m_radiuses, model);
            dump_neighbours(grid, field, p, prop, 1, counts, idx_out, max_nb);
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// The reference search template (calc_cov_field with the unclipped model).
int ref_template(int type, const double* ranges, const double* angles, double sill, double nugget,
                 const int* radiuses, int* out, int cap) {
    try {
        hpgl::ok_params_t p;
        fill_cov(p, type, ranges, angles, sill, nugget);
        p.set_radiuses(radiuses[0], radiuses[1], radiuses[2]);
        hpgl::cov_model_t model(p);
        std::vector<hpgl::sugarbox_vector_t> v;
        hpgl::calc_cov_field<hpgl::cov_model_t, hpgl::sugarbox_location_t>(p.m_radiuses, model, v);
        for (size_t i = 0; i < v.size() && (int)i < cap; ++i) {
            out[3 * i + 0] = v[i][0];
            out[3 * i + 1] = v[i][1];
            out[3 * i + 2] = v[i][2];
        }
        return (int)v.size();
    } catch (const std::exception&) {
        return -1;
    }
}

// hpgl::simple_kriging_weights (the path behind hpgl_simple_kriging_weights).
int ref_sk_weights(int type, const double* ranges, const double* angles, double sill, double nugget,
                   const float* center, const float* xs, const float* ys, const float* zs, int n, double* w,
                   double* variance) {
    try {
        hpgl::sk_params_t p;
        fill_cov(p, type, ranges, angles, sill, nugget);
        hpgl::real_location_t c(center[0], center[1], center[2]);
        std::vector<hpgl::real_location_t> coords(n);
        for (int i = 0; i < n; ++i) {
            coords[i][0] = xs[i];
            coords[i][1] = ys[i];
            coords[i][2] = zs[i];
        }
        std::vector<hpgl::kriging_weight_t> weights;
        double var = -1;
        hpgl::simple_kriging_weights(&p, c, coords, weights, var);
        for (int i = 0; i < n; ++i) w[i] = i < (int)weights.size() ? weights[i] : 0.0;
        if (variance) *variance = var;
        return (int)weights.size();
    } catch (const std::exception&) {
        return -1;
    }
}

}  // extern "C"
