/* hpgl.h -- the reference's kriging C ABI, served by the MI355X library.
 *
 * libhpgl.so exports these symbols with the reference's names, struct
 * layouts and calling convention, so the reference's ctypes binding
 * (src/geo_bsd/hpgl_wrap.py) and geo.py load it unchanged.  Each entry point
 * takes host buffers, copies them to the GPU, runs the HIP pipeline and copies
 * the result back before returning.
 *
 * Citations are to the reference repository (itohnobue/hpgl-reborn).
 */
#ifndef HPGL_MI355X_HPGL_H
#define HPGL_MI355X_HPGL_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/geo_bsd/hpgl/api.h:18-22 */
struct hpgl_shape_t {
    int m_data[3];
    int m_strides[3];
};

/* src/geo_bsd/hpgl/api.h:24-29 */
struct hpgl_cont_masked_array_t {
    float* m_data;
    unsigned char* m_mask;
    struct hpgl_shape_t m_shape;
};

/* src/geo_bsd/hpgl/api.h:31-37 */
struct hpgl_ind_masked_array_t {
    unsigned char* m_data;
    unsigned char* m_mask;
    struct hpgl_shape_t m_shape;
    int m_indicator_count;
};

/* src/geo_bsd/hpgl/api.h:51-60 */
struct hpgl_ok_params_t {
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
};

/* src/geo_bsd/hpgl/api.h:62-73 */
struct hpgl_sk_params_t {
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    unsigned char m_automatic_mean;
    double m_mean;
};

/* src/geo_bsd/hpgl/api.h:75-82 */
struct hpgl_cov_params_t {
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
};

/* src/geo_bsd/hpgl/api.h:98-108 */
struct hpgl_median_ik_params_t {
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    double m_marginal_probs[2];
};

/* src/geo_bsd/hpgl/api.h:110-120 (one per category of indicator kriging) */
struct hpgl_ik_params_t {
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    double m_marginal_prob;
};

/* src/geo_bsd/hpgl/api.h:39-43 */
struct hpgl_float_array_t {
    float* m_data;
    struct hpgl_shape_t m_shape;
};

/* src/geo_bsd/hpgl/api.h:45-49 */
struct hpgl_ubyte_array_t {
    unsigned char* m_data;
    struct hpgl_shape_t m_shape;
};

/* src/geo_bsd/hpgl/api.h:84-96; m_kriging_kind 0 = ordinary, 1 = simple */
struct hpgl_sgs_params_t {
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    int m_kriging_kind;
    long m_seed;
    int m_min_neighbours;
};

/* src/geo_bsd/hpgl/api.h:148-153: ascending values and their cumulative
 * probabilities */
struct hpgl_non_parametric_cdf_t {
    float* m_values;
    float* m_probs;
    long long m_size;
};

/* Replaces api.cpp:38-41.  Message of the last failed call ("" if none). */
char* hpgl_get_last_exception_message(void);

/* Replace output.cpp:58-68.  Messages and progress of the kriging calls are
 * routed to these callbacks; without them messages go to stdout prefixed
 * "[LOG2]" as in the reference. */
void hpgl_set_output_handler(int (*handler)(char* str, void* param), void* param);
void hpgl_set_progress_handler(int (*handler)(char* stage, int percentage, void* param), void* param);

/* Replace api.cpp:43-51.  Kept for ABI compatibility: the value is stored and
 * returned; the GPU path does not use host threads. */
void hpgl_set_thread_num(int n_threads);
int hpgl_get_thread_num(void);

/* Replaces api.cpp:253-293 (geo.ordinary_kriging).  Output buffers hold a
 * clone of the input on entry; kriged cells get the estimate and mask 1,
 * cells that cannot be kriged keep the clone. */
void hpgl_ordinary_kriging(struct hpgl_cont_masked_array_t* input_data, struct hpgl_ok_params_t* params,
                           struct hpgl_cont_masked_array_t* output_data);

/* Replaces api.cpp:323-354 (geo.simple_kriging).  Cells that cannot be
 * kriged get the mean. */
void hpgl_simple_kriging(float* input_data, unsigned char* input_mask, struct hpgl_shape_t* input_data_shape,
                         struct hpgl_sk_params_t* params, float* output_data, unsigned char* output_mask,
                         struct hpgl_shape_t* output_data_shape);

/* Replaces api.cpp:356-410 (geo.simple_kriging_weights): SK weights for
 * float coordinates with the unclipped model.  Returns 0, or -1 on error. */
int hpgl_simple_kriging_weights(float* center_coords, float* neighbours_x, float* neighbours_y, float* neighbours_z,
                                int neighbours_count, struct hpgl_cov_params_t* params, float* weights);

/* Replaces api.cpp:552-601 (geo.median_ik): every cell gets indicator 0/1
 * (1 when the kriged P(data == 1) exceeds 0.5) and mask 1. */
void hpgl_median_ik(struct hpgl_ind_masked_array_t* in_data, struct hpgl_median_ik_params_t* params,
                    struct hpgl_ind_masked_array_t* out_data);

/* Replaces api.cpp:485-516 (geo_bsd.sgs_simulation): sequential Gaussian
 * simulation in place on `data` along the reference's random path with its
 * draws (seed), simple kriging with mean 0 or ordinary kriging over the
 * covariance of the radius box, an optional normal-score transform (cdf, may
 * be null) and an optional mask of the cells to simulate (== 1; may be null).
 * `mean` is accepted and unused, as in the reference.  Errors: see
 * hpgl_get_last_exception_message(). */
void hpgl_sgs_simulation(struct hpgl_cont_masked_array_t* data, struct hpgl_sgs_params_t* params,
                         struct hpgl_non_parametric_cdf_t* cdf, double* mean,
                         struct hpgl_ubyte_array_t* simulation_mask);

/* Replaces api.cpp:519-548 (geo_bsd.sgs_simulation with a mean array):
 * simple kriging with one mean per cell (taken through the normal-score
 * transform with the data when cdf is given). */
void hpgl_sgs_lvm_simulation(struct hpgl_cont_masked_array_t* data, struct hpgl_sgs_params_t* params,
                             struct hpgl_non_parametric_cdf_t* cdf, struct hpgl_float_array_t* means,
                             struct hpgl_ubyte_array_t* simulation_mask);

/* Replaces api.cpp:412-455 (geo.lvm_kriging, lvm_kriging.cpp:16-39): simple
 * kriging where every cell has its own mean, mean_data (N floats, the
 * input's shape).  Cells that cannot be kriged get their mean. */
void hpgl_lvm_kriging(float* input_data, unsigned char* input_mask, struct hpgl_shape_t* input_data_shape,
                      float* mean_data, struct hpgl_shape_t* mean_data_shape, struct hpgl_ok_params_t* params,
                      float* output_data, unsigned char* output_mask, struct hpgl_shape_t* output_data_shape);

/* Replaces api.cpp:457-482 (geo.indicator_kriging with more than two
 * categories, indicator_kriging.h:88-161): per category c, simple kriging of
 * (data == c) with mean params[c].m_marginal_prob and that category's model,
 * radiuses and neighbour limit; the probabilities get the order-relation
 * correction and every cell gets the first most probable category, mask 1. */
void hpgl_indicator_kriging(struct hpgl_ind_masked_array_t* in_data, struct hpgl_ind_masked_array_t* out_data,
                            struct hpgl_ik_params_t* params, int indicator_count);

/* Property files.  The file is read on the host and its text parsed on the
 * GPU (or printed on the GPU and written on the host); values and masks are
 * host arrays as in the reference.  All return 0, or -1 with the message in
 * hpgl_get_last_exception_message(). */

/* Replaces api.cpp:53-75 (read_inc_file.cpp:137-165): the first line that
 * starts with a letter names the property; then `size` values as
 * sscanf("%f") reads them ("--" comments to the end of the line, '/' ends the
 * data); mask = value != undefined_value. */
int hpgl_read_inc_file_float(char* filename, float undefined_value, int size, float* data, unsigned char* mask);

/* Replaces api.cpp:77-116 (read_inc_file.cpp:167-193): integers as
 * sscanf("%d"), stored as bytes, mask = value != undefined_value; informed
 * bytes equal to values[j] become j. */
int hpgl_read_inc_file_byte(char* filename, int undefined_value, int size, unsigned char* data, unsigned char* mask,
                            unsigned char* values, int values_count);

/* Replaces api.cpp:118-141 (property_writer.cpp:44-63): the name line, one
 * "%E" line per value (undefined_value where the mask is 0), then "/". */
int hpgl_write_inc_file_float(char* filename, struct hpgl_cont_masked_array_t* arr, float undefined_value,
                              char* name);

/* Replaces api.cpp:143-198 (property_writer.cpp:65-85): "%d" lines of
 * (signed char) values[data] for informed cells. */
int hpgl_write_inc_file_byte(char* filename, struct hpgl_ind_masked_array_t* arr, int undefined_value, char* name,
                             unsigned char* values, int values_count);

/* Replace api.cpp:200-251 (property_writer.cpp:88-139): GSLIB files, header
 * "HPGL saved GSLIB file", 1, the name line, then the values. */
int hpgl_write_gslib_cont_property(struct hpgl_cont_masked_array_t* data, const char* filename, const char* name,
                                   double undefined_value);
int hpgl_write_gslib_byte_property(struct hpgl_ind_masked_array_t* data, const char* filename, const char* name,
                                   double undefined_value, unsigned char* values, int values_count);

#ifdef __cplusplus
}
#endif

#endif /* HPGL_MI355X_HPGL_H */
