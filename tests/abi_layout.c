/* Prints the size of every struct include/flashraw_hip.h declares and the offset and size of each
 * of its fields, one line each: "<struct> <field|-> <offset> <size>". The CPU test
 * tests/test_abi_layout.py compiles this with gcc and checks the ctypes mirror in
 * flashraw_amd/gpu.py (and the Rust #[repr(C)] mirror's field order) against it. */
#include <stddef.h>
#include <stdio.h>

#include "flashraw_hip.h"

#define S(T) printf(#T " - 0 %zu\n", sizeof(T));
#define F(T, f) printf(#T " " #f " %zu %zu\n", offsetof(T, f), sizeof(((T *)0)->f));

int main(void) {
    S(frh_rect) F(frh_rect, x) F(frh_rect, y) F(frh_rect, width) F(frh_rect, height)

    S(frh_frame) F(frh_frame, size) F(frh_frame, input) F(frh_frame, data) F(frh_frame, data_len)
    F(frh_frame, data_on_device) F(frh_frame, raster_width) F(frh_frame, raster_height) F(frh_frame, cpp)
    F(frh_frame, window) F(frh_frame, origin_x) F(frh_frame, origin_y) F(frh_frame, linearization)
    F(frh_frame, linearization_len) F(frh_frame, bits) F(frh_frame, order) F(frh_frame, row_stride)

    S(frh_frame_info) F(frh_frame_info, size) F(frh_frame_info, cfa_width) F(frh_frame_info, cfa_height)
    F(frh_frame_info, cfa) F(frh_frame_info, black_rows) F(frh_frame_info, black_cols)
    F(frh_frame_info, black_samples) F(frh_frame_info, black_values) F(frh_frame_info, black_values_len)
    F(frh_frame_info, black_row_deltas) F(frh_frame_info, black_row_deltas_len)
    F(frh_frame_info, black_col_deltas) F(frh_frame_info, black_col_deltas_len)
    F(frh_frame_info, white_values) F(frh_frame_info, white_values_len)

    S(frh_edge_spec) F(frh_edge_spec, size) F(frh_edge_spec, period) F(frh_edge_spec, cx) F(frh_edge_spec, cy)
    F(frh_edge_spec, beta) F(frh_edge_spec, halo) F(frh_edge_spec, canon) F(frh_edge_spec, canon_len)
    F(frh_edge_spec, g0) F(frh_edge_spec, g0_len) F(frh_edge_spec, dir) F(frh_edge_spec, dir_len)
    F(frh_edge_spec, ahd)

    S(frh_curve) F(frh_curve, size) F(frh_curve, lut8) F(frh_curve, lut8_len) F(frh_curve, lut16)
    F(frh_curve, lut16_len)

    S(frh_tone) F(frh_tone, size) F(frh_tone, table) F(frh_tone, table_len) F(frh_tone, xmax)
    F(frh_tone, scale) F(frh_tone, to_prophoto) F(frh_tone, white)

    S(frh_develop_spec) F(frh_develop_spec, size) F(frh_develop_spec, crop) F(frh_develop_spec, half_size)
    F(frh_develop_spec, wb) F(frh_develop_spec, cam_to_rgb) F(frh_develop_spec, orientation)
    F(frh_develop_spec, output) F(frh_develop_spec, output2) F(frh_develop_spec, has_levels)
    F(frh_develop_spec, levels) F(frh_develop_spec, edge) F(frh_develop_spec, curve) F(frh_develop_spec, tone)
    F(frh_develop_spec, cancel)

    S(frh_timings) F(frh_timings, upload_ms) F(frh_timings, compute_ms) F(frh_timings, kernel_ms)
    F(frh_timings, has_kernel_ms) F(frh_timings, bands)

    S(frh_dn_spec) F(frh_dn_spec, size) F(frh_dn_spec, width) F(frh_dn_spec, height) F(frh_dn_spec, stride)
    F(frh_dn_spec, black) F(frh_dn_spec, sig_a) F(frh_dn_spec, sig_b) F(frh_dn_spec, threshold)
    F(frh_dn_spec, step) F(frh_dn_spec, precision) F(frh_dn_spec, seg_rows)

    S(frh_nr_plane) F(frh_nr_plane, levels) F(frh_nr_plane, gain) F(frh_nr_plane, steps)
    F(frh_nr_plane, hard_only0) F(frh_nr_plane, keep)

    S(frh_nr_spec) F(frh_nr_spec, size) F(frh_nr_spec, has_luma) F(frh_nr_spec, has_chroma) F(frh_nr_spec, luma)
    F(frh_nr_spec, chroma) F(frh_nr_spec, chroma_radius) F(frh_nr_spec, hard) F(frh_nr_spec, max_blocks)
    F(frh_nr_spec, output) F(frh_nr_spec, orientation) F(frh_nr_spec, tone) F(frh_nr_spec, curve)
    F(frh_nr_spec, cancel)

    S(frh_server_config) F(frh_server_config, size) F(frh_server_config, device_budget_bytes)
    F(frh_server_config, max_batch_frames) F(frh_server_config, max_inflight_batches)

    S(frh_job_result) F(frh_job_result, size) F(frh_job_result, status) F(frh_job_result, queue_ms)
    F(frh_job_result, upload_ms) F(frh_job_result, batch_wait_ms) F(frh_job_result, latency_ms)
    F(frh_job_result, batch_frames)

    S(frh_server_stats) F(frh_server_stats, size) F(frh_server_stats, jobs_submitted)
    F(frh_server_stats, jobs_done) F(frh_server_stats, jobs_failed) F(frh_server_stats, batches)
    F(frh_server_stats, batched_frames) F(frh_server_stats, peak_device_bytes)
    F(frh_server_stats, device_budget_bytes)
    return 0;
}
