"""Parity status of every oracle function: which check pins it.

TEST INFRASTRUCTURE ONLY.  A function is "pinned" when a check that does not
restate its formula would fail on a plausible mistake in it (a dropped term,
a wrong sign or index, a transposed operand): a published known-answer test,
a closed form from the reference's own tests, brute force on tiny inputs, or
an invariant.  "parity unpinned" marks what only an import of the reference
(or of a closed library it calls) could pin.

tests/test_oracle_pinning.py asserts that every public oracle function has an
entry here.
"""

PINS = {
    # philox.py
    "philox.philox4x32_10": ("pinned", "Random123 Philox4x32-10 known-answer vectors "
                             "(test_philox_known_answers)"),
    "philox.draw": ("pinned", "stream layout through philox4x32_10 KATs; GPU bit-exact "
                    "(test_device_exp_and_philox_bit_exact)"),
    "philox.u01f": ("pinned", "range (0,1) and exact dyadic values (test_uniform_ranges_and_exactness)"),
    "philox.u01d": ("pinned", "range (0,1] and 53-bit exactness (test_uniform_ranges_and_exactness)"),
    "philox.normal4": ("pinned", "moments of 4e5 draws (test_normals_are_standard)"),
    "philox.predict_normals": ("pinned", "moments and independence across components, group "
                               "slots and groups (test_predict_normals_moments_and_independence); "
                               "velocity-noise moments, RBPF noise covariance; GPU particle "
                               "clouds equal the oracle's"),
    "philox.xo_seed": ("pinned", "Philox KATs for the seed block; zero-state rule "
                       "(test_xoshiro_outputs_and_seed_rules); GPU bit-exact (gg_debug_math op 3)"),
    "philox.xo_next": ("pinned", "state transition T with T^(2^64) equal to the published jump "
                       "polynomial of xoshiro128+ (test_xoshiro_transition_matches_published_"
                       "jump_polynomial); hand-computed first outputs; uniformity and lag-1 "
                       "correlation; GPU bit-exact"),
    "philox.xo_normals": ("pinned", "Box-Muller of xo_next uniforms, moments through "
                          "test_predict_normals_moments_and_independence"),
    # detmath.py
    "detmath.tile_scale": ("pinned", "c 2^k = exp(m) against libm, c in [1, 2), edge cases "
                           "(test_tile_scale_and_shift_exact)"),
    "detmath.tile_shift": ("pinned", "exact against rational arithmetic; 0 below 2^-1000 "
                           "(test_tile_scale_and_shift_exact)"),
    "detmath.exp64": ("pinned", "<= 2 ulp against libm over [-745, 0] (test_exp64_against_libm)"),
    "detmath.exp32": ("pinned", "<= 4 ulp (binary32) against binary64 exp "
                      "(test_exp32_against_float64_exp)"),
    "detmath.fmaf": ("pinned", "bit-exact against the correctly rounded binary32 of the exact "
                     "rational a*b + c, ties included (test_fmaf_against_rational_arithmetic)"),
    "detmath.block_scan256": ("pinned", "exact on integers, rational bound on reals "
                              "(test_block_scan_exact_on_integers, test_block_scan_close_to_exact_rational)"),
    "detmath._hs_wave": ("pinned", "inside block_scan256 (exact on integers)"),
    "detmath.block_total256": ("pinned", "last element of the pinned scan"),
    "detmath.tile_view": ("pinned", "layout only; exercised by the brute-force systematic tests"),
    "detmath.tile_partials": ("pinned", "brute-force systematic + ESS closed forms"),
    "detmath.kcount": ("pinned", "searchsorted over the exactly rounded targets; brute force "
                       "through the systematic tests"),
    "detmath.decide": ("pinned", "brute-force systematic (test_systematic_matches_brute_force)"),
    "detmath.systematic_targets": ("pinned", "pf_device.cu:1202 (u0+j)/N; brute force"),
    "detmath.reference_u0": ("pinned", "int32 product of pf_device.cu:1852 on N where 1e6*N "
                             "overflows (test_reference_u0_int32_product)"),
    "detmath.fresh_u0": ("pinned", "range [0,1) and Philox stream"),
    "detmath.tile_ends": ("pinned", "K_0 = 0 and K_nt = N, non-decreasing, equal to the "
                          "brute-force count of targets below each tile's share of the binary64 "
                          "cumulative weight (test_tile_ends_pins)"),
    "detmath.systematic_ends": ("pinned", "brute force cumsum+searchsorted, invariants "
                                "(test_systematic_matches_brute_force, test_systematic_invariants_large)"),
    "detmath.systematic_ancestors": ("pinned", "brute force on tiny N, identity for uniform weights, "
                                     "single heavy particle"),
    "detmath.megopolis_ref_ancestors": ("pinned", "B=0 identity, uniform weights accept every "
                                        "proposal, slot-weight rule of pf_device.cu:1262-1270 "
                                        "on a hand-computed case"),
    "detmath.megopolis_ancestors": ("pinned", "chi-square against the target distribution "
                                    "(test_megopolis_chesser_matches_target_distribution); "
                                    "acceptance probability min(1, w_j/w_k) measured "
                                    "(test_megopolis_acceptance_rule)"),
    "detmath.megopolis_weights": ("pinned", "max weight 1, ratios exp(lw_j - lw_k) to binary32 "
                                  "(test_megopolis_acceptance_rule)"),
    "detmath.sequential_systematic": ("pinned", "textbook searchsorted(side='left') on tiny N"),
    # pf.py
    "pf.enu_rotation": ("pinned", "orthonormality and the up axis along the ellipsoid normal "
                        "(test_scenario_geometry)"),
    "pf.huber_cost": ("pinned", "closed form -28.875 of test_cuda_streams.py:689-692"),
    "pf.log_likelihood": ("pinned", "closed forms -200.0 (Gaussian) and -28.875 (Huber); "
                          "zero residual -> 0; gate all-pass / all-reject; Student-t brute force; "
                          "inter-system biases: 0 at the truth, -n_k delta^2 / (2 sigma^2) for "
                          "ISB k-1 off by delta (test_isb_closed_forms)"),
    "pf.log_likelihood_gmm": ("pinned", "the two-component normal mixture density from "
                              "scipy.stats.norm (unit satellite weights, up to the dropped "
                              "-log sqrt(2 pi)); w_los = 1 -> Gaussian; equal components -> "
                              "independent of w_los (test_gmm_log_likelihood_pins)"),
    "pf.ess": ("pinned", "uniform -> N, one dominant -> < 2 (test_particle_filter.py:174-186)"),
    "pf.weighted_mean": ("pinned", "uniform weights -> arithmetic mean (test_particle_filter.py:293-315)"),
    "pf.position_spread": ("pinned", "uniform weights -> RMS distance, closed form on a 2-point cloud"),
    "pf.OraclePF.initialize": ("pinned", "moments; the draw order x,y,z,cb follows pf_device.cu:67-104. "
                               "The reference's cuRAND stream itself is parity unpinned"),
    "pf.OraclePF.predict": ("pinned", "sigma=0 closed form p += v*dt (test_cuda_streams.py:239-255); "
                            "guide blend and alpha clamp closed forms; velocity-noise moments; "
                            "RBPF Sigma_v 4.5 / 1.5 and position std (test_cuda_streams.py:257-283, "
                            "486-505); RBPF noise covariance sigma^2 I + dt^2 Sigma_v for a "
                            "non-diagonal Sigma_v; noise values parity unpinned (cuRAND)"),
    "pf.OraclePF.weight": ("pinned", "as log_likelihood"),
    "pf.OraclePF.resample_systematic": ("pinned", "as systematic_ancestors"),
    "pf._gated_sum": ("pinned", "inside the DD closed forms (gate: all kept / too few kept)"),
    "pf.dd_pseudorange_loglik": ("pinned", "reference closed forms -5000 / -148.875 "
                                 "(test_cuda_streams.py:696-737; test_dd_closed_forms)"),
    "pf.afv": ("pinned", "integer cycles map to 0, x+k to x, |afv| <= 1/2 (test_dd_closed_forms)"),
    "pf.carrier_afv_loglik": ("pinned", "zero at the truth, integer-cycle invariance, a 0.3-cycle "
                              "offset gives -0.5 (0.3/sigma)^2 (test_dd_closed_forms)"),
    "pf.dd_carrier_afv_loglik": ("pinned", "reference closed forms -40.5 / -12.375 "
                                 "(test_cuda_streams.py:739-771; test_dd_closed_forms)"),
    "pf._los": ("pinned", "inside the Doppler checks"),
    "pf.doppler_update": ("pinned", "noise-free rows: gain 1 returns the true velocity, "
                          "zero misfit at the truth for any clock drift, the step cap "
                          "(test_doppler_oracle_known_answers; test_cuda_streams.py:285-309)"),
    "pf.doppler_kf_update": ("pinned", "a vague prior converges to the true velocity and the "
                             "covariance shrinks; one row against the closed-form scalar update "
                             "(test_doppler_oracle_known_answers; test_cuda_streams.py:311-335)"),
    "pf.OraclePF.resample_megopolis": ("pinned", "as megopolis_ref_ancestors"),
    "pf.OraclePF.resample_megopolis_chesser": ("pinned", "as megopolis_ancestors"),
    # wls.py
    "wls.wls_solve": ("pinned", "noise-free truth < 1 cm with and without the Sagnac term; "
                      "singular geometry; < 4 satellites"),
    "wls.wls_position_ref": ("pinned", "test_verification.py:305-333 truth (no Sagnac)"),
    "wls.wls_batch_ref": ("pinned", "test_positioning.py:49-75 truth (Sagnac)"),
    "wls._los": ("pinned", "Sagnac closed form of range_model.py:20-29 (test_sagnac_rotation_closed_form)"),
    "wls._solve4": ("pinned", "numpy.linalg.solve on random SPD systems; singular pivot rule"),
    "wls._solve_n": ("pinned", "numpy.linalg.solve on random SPD systems n = 3..8; a zero "
                     "row/column gives 0 for its unknown and the reduced solve for the rest "
                     "(test_solve_n_against_numpy_and_zero_rows)"),
    "wls.wls_multi_gnss_solve": ("pinned", "test_multi_gnss.py:109-151 ISB recoveries (Sagnac); "
                                 "one system equals the pinned 4-state WLS; the noisy optimum "
                                 "equals scipy least_squares; empty system and foreign id "
                                 "(test_multi_gnss_*)"),
    "wls.wls_multi_gnss_ref": ("pinned", "plain ranges recovered, Sagnac ranges not "
                               "(multi_gnss.cu:101-112); guards of test_multi_gnss.py:188-197"),
    "wls.wls_multi_gnss_batch_ref": ("pinned", "test_multi_gnss.py:100-172 truths and bounds"),
    # raytrace.py
    "raytrace.mt_hit": ("pinned", "plane parameter and barycentric 3x3 solve on random rays "
                        "(test_mt_hit_against_plane_and_barycentric_solve)"),
    "raytrace.los_blocked": ("pinned", "box cases of test_bvh.py:26-39, empty mesh"),
    "raytrace.los_check": ("pinned", "test_bvh.py:26-39, 178-191 (NaN padding)"),
    "raytrace.multipath": ("pinned", "wall reflection closed form "
                           "(test_multipath_wall_closed_form)"),
    "raytrace.mixture": ("pinned", "log((1-p)e^los + p e^nlos) closed form"),
    "raytrace.weight_3d": ("pinned", "closed forms of test_pf3d_bvh.py:94-194"),
    # ffbsi.py
    "ffbsi.choice_index": ("pinned", "numpy Generator.choice on 40 seeds "
                           "(test_choice_index_is_generator_choice)"),
    "ffbsi.transition_logpdf": ("pinned", "scipy.stats.norm.logpdf sums up to a constant"),
    "ffbsi._normalised": ("pinned", "through ffbsi_indices"),
    "ffbsi.ffbsi_indices": ("pinned", "test_ffbsi.py:37-50 bounds; chi-square of the backward "
                            "draw against w0 p(x1 | x0); generator consumption incl. "
                            "degenerate steps"),
    "ffbsi.genealogy_indices": ("pinned", "test_ffbsi.py:23-35 trace invariant"),
    # learned.py
    "learned.unpack": ("pinned", "inverse of pack (round trip in the GPU parameter test)"),
    "learned.pack": ("pinned", "inverse of unpack"),
    "learned.init_params": ("pinned", "shapes and the b3 = [0, ln 25] start; no numerics"),
    "learned.bf16": ("pinned", "round-to-nearest-even on hand-checked bit patterns "
                     "(test_bf16_rounding_known_values)"),
    "learned.q_e4m3": ("pinned", "every finite OCP E4M3FN value is a fixed point, midpoints go "
                       "to the even neighbour, saturation at 448, subnormal quantum 2^-9 "
                       "(test_e4m3_quantiser_known_values)"),
    "learned.e4m3_scale": ("pinned", "smallest power of two with max|W|/s <= 448 "
                           "(test_e4m3_quantiser_known_values)"),
    "learned.quantised": ("pinned", "fp8 mode applies q_e4m3 with e4m3_scale per tensor; "
                          "other modes return the parameters (test_e4m3_quantiser_known_values)"),
    "learned.first_layer_bias": ("pinned", "binary32 left-to-right sum against a binary64 sum"),
    "learned.scaled_residual": ("pinned", "clamp bounds"),
    "learned.mlp_forward": ("pinned", "exact mode: a ReLU network is affine on a fixed activation "
                            "pattern (test_mlp_forward_affine_on_fixed_pattern)"),
    "learned.pair_loglik": ("pinned", "closed form of a Gaussian log-density"),
    "learned.mlp_backward": ("pinned", "central finite differences in binary64 "
                             "(test_contrastive_gradient_finite_differences)"),
    "learned.candidate_residuals": ("pinned", "u.d - dcb against binary64"),
    "learned.contrastive_loss_and_grad": ("pinned", "finite differences; loss >= 0 and = 0 "
                                          "for a perfect softmax match"),
}

UNPINNED = {
    "cuRAND normal/uniform streams (pf_device.cu:118-186,1245-1270)":
        "closed NVIDIA library; the library uses its own Philox4x32-10 and the oracle its NumPy twin",
    "thrust::inclusive_scan summation order (pf_device.cu:1847-1849)":
        "unspecified by thrust; the library pins its own fixed-order scan",
}
