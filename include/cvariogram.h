/* cvariogram.h -- the experimental-variogram ABI of the reference's
 * _cvariogram library (src/geo_bsd/_cvariogram/api.h:24-144), exported by
 * libhpgl.so and computed on the GPU.  Struct layouts are the reference's,
 * as its cvariogram.py declares them (cvariogram.py:38-86).
 */
#ifndef HPGL_MI355X_CVARIOGRAM_H
#define HPGL_MI355X_CVARIOGRAM_H

#ifdef __cplusplus
extern "C" {
#endif

struct vector_t {
    double m_data[3];
};

struct ellipsoid_t {
    struct vector_t m_direction1;
    struct vector_t m_direction2;
    struct vector_t m_direction3;
    double m_R1, m_R2, m_R3;
};

struct variogram_search_template_t {
    double m_lag_width;
    double m_lag_separation;
    double m_tol_distance;
    int m_num_lags;
    double m_first_lag_distance;
    struct ellipsoid_t m_ellipsoid;
};

struct cont_point_set_t {
    float* xs;
    float* ys;
    float* zs;
    float* values;
    int size;
};

struct hard_data_t {
    float* m_data;
    unsigned char* m_mask;
    int m_data_shape[3];
    int m_data_strides[3];
    int m_mask_shape[3];
    int m_mask_strides[3];
};

/* Replaces ellipsoid.cpp:37-72: the three direction columns of
 * Rz(azimuth) Ry(dip) Rx(rotation), angles in degrees. */
void fill_ellipsoid_directions(struct ellipsoid_t* ell, double azimuth, double dip, double rotation);

/* Replaces variograms.cpp:353-448: semivariogram of gridded hard data
 * (strides in elements) over the first min(num_lags, result_length) lags;
 * 0 for a lag without pairs.  percentToUse < 100 keeps a deterministic
 * pseudo-random share of the pairs (the reference draws rand()). */
void calc_variograms(struct variogram_search_template_t* templ, struct hard_data_t* data, float* result_covariations,
                     int result_length, int percentToUse);

/* Replaces variograms.cpp:450-506: semivariogram of a point set over every
 * ordered pair of points (self pairs included, as in the reference). */
void calc_variograms_from_point_set(struct variogram_search_template_t* templ, struct cont_point_set_t* point_set,
                                    float* result_covariations, int result_length);

/* Device forms: data, mask, xyzv (x, y, z, value per point) and the result
 * are device arrays; the grid is dense, x fastest.  Return 0, or -1 with
 * hpglx_last_error() set.  Both synchronise `stream`. */
int hpglx_calc_variograms(const struct variogram_search_template_t* templ, const float* d_data,
                          const unsigned char* d_mask, const int dims[3], int lag_count, int percent, float* d_result,
                          void* stream);
int hpglx_calc_variograms_points(const struct variogram_search_template_t* templ, const float* d_xyzv, int n,
                                 int lag_count, float* d_result, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HPGL_MI355X_CVARIOGRAM_H */
