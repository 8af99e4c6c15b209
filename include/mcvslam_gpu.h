/*
 * mcvslam_gpu.h — C ABI of the MI355X (gfx950) hot path of MCVSLAM.
 *
 * Every entry point returns MCV_OK (0) or a negative MCV_ERR_* code; the
 * message of the last failure on a context is mcv_last_error(ctx).  No C++
 * exception crosses this boundary.  Buffers are owned by the caller; host
 * pointers are ordinary memory, "_dev" entry points take device pointers
 * (HBM) and enqueue on the context's stream without synchronizing.
 * Contexts are independent: one per thread replaces the reference's one
 * ORB extractor per camera thread (src/Frame.cpp:24-26, :120-125).
 */
#ifndef MCVSLAM_GPU_H
#define MCVSLAM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCV_OK 0
#define MCV_ERR_INVALID (-1)   /* bad argument */
#define MCV_ERR_HIP (-2)       /* HIP runtime failure */
#define MCV_ERR_CAPACITY (-3)  /* output too small; required size reported */
#define MCV_ERR_EMPTY (-4)     /* empty image (ORBextractor::operator() returns -1) */
#define MCV_ERR_STATE (-5)     /* not configured, or device work overflowed */

#define MCV_FLAG_GRAPH 1       /* replay the launch sequence as a captured hipGraph */
#define MCV_FLAG_CHECK 2       /* wait for the run and report device-side overflow (see below) */

typedef struct mcv_ctx mcv_ctx;
typedef struct mcv_rig mcv_rig;
typedef struct mcv_ba mcv_ba;

/* cv::KeyPoint without class_id: pt.x, pt.y, size, angle (deg), response, octave */
typedef struct {
    float x, y, size, angle, response;
    int32_t octave;
} mcv_keypoint;

/* ---- context --------------------------------------------------------- */
int mcv_create(mcv_ctx** ctx, int device);
void mcv_destroy(mcv_ctx* ctx);
const char* mcv_last_error(const mcv_ctx* ctx);
/* hipStream_t to enqueue on; NULL is the null (default) stream.  A new
 * context enqueues on a non-blocking stream of its own. */
int mcv_set_stream(mcv_ctx* ctx, void* hip_stream);
int mcv_use_own_stream(mcv_ctx* ctx);
int mcv_synchronize(mcv_ctx* ctx);
/* Synchronous device -> host copy of `bytes` on the context's stream. */
int mcv_copy_to_host(mcv_ctx* ctx, void* dst, const void* d_src, uint64_t bytes);

/* ---- ORB extraction ---------------------------------------------------
 * Replaces ORB_SLAM3::ORBextractor::init (modules/local_feature/ORB/
 * orb3_extractor/ORBextractor.cc:407-457) as called by ORB::ORB / ORB::Parse
 * (modules/local_feature/ORB/ORBExtractor.cpp:10-23). */
int mcv_orb_configure(mcv_ctx* ctx, int nfeatures, float scale_factor, int nlevels, int ini_th_fast,
                      int min_th_fast);
/* mvScaleFactor, mvInvScaleFactor, mvLevelSigma2, mvInvLevelSigma2,
 * mnFeaturesPerLevel (ORBextractor.h:96-99); each array has nlevels entries;
 * any pointer may be NULL */
int mcv_orb_tables(const mcv_ctx* ctx, float* scale, float* inv_scale, float* sigma2, float* inv_sigma2,
                   int32_t* n_per_level);
/* Upper bound of the keypoints one image of this size can produce. */
int mcv_orb_capacity(const mcv_ctx* ctx, int width, int height, int n_seeds, int32_t* capacity);
/* Replaces ORBextractor::operator()(image, mask, keypoints, descriptors, lap)
 * (ORBextractor.h:57-58, ORBextractor.cc:831-899) behind BaseExtractor::Extract
 * (modules/local_feature/BaseExtractor/BaseExtractor.hpp:18, ORBExtractor.cpp:27-38).
 * image: width x height u8 with row stride `stride` (host).  seeds: keypoints
 * appended to their octave after the octree (the reference's input
 * `_keypoints`), may be NULL.  Output: `count` keypoints and count x 32
 * descriptor bytes, in the reference's order.  Returns MCV_ERR_EMPTY for an
 * empty image, MCV_ERR_CAPACITY (count = required) when capacity is short. */
int mcv_orb_extract(mcv_ctx* ctx, const uint8_t* image, int width, int height, int stride, const mcv_keypoint* seeds,
                    int n_seeds, mcv_keypoint* kps, uint8_t* desc, int capacity, int32_t* count);
/* mvImagePyramid[level] of the last mcv_orb_extract (ORBextractor.h:72, read
 * by Frame::ComputeStereoMatch, src/Frame.cpp:239-261). */
int mcv_orb_pyramid(mcv_ctx* ctx, int level, uint8_t* out, int capacity, int32_t* width, int32_t* height);
/* Batched extraction of device-resident images in one launch sequence: the
 * reference's per-camera ORBextractor::operator() calls
 * (ORBextractor.cc:831-899, src/Frame.cpp:118-125) for n images at once.
 * d_images: host array of n device pointers; widths/heights/strides: host
 * arrays.  seeds/n_seeds (both may be NULL): per image a host array of
 * n_seeds[i] keypoints appended to their octave after the octree, as the
 * reference's input `_keypoints` (ORBextractor.cc:845-847; the KL-tracked
 * points of src/Frame.cpp:34-74).  Outputs (device): d_kps[n * capacity],
 * d_desc[n * capacity * 32], d_counts[n].  capacity must be >=
 * mcv_orb_capacity(width, height, n_seeds) of every image.
 * flags: MCV_FLAG_GRAPH replays a captured graph (recaptured whenever an
 * output pointer or the capacity changes); MCV_FLAG_CHECK waits for the run
 * and, when the device-side octree node pool overflowed, grows it and runs
 * again (MCV_ERR_STATE after three attempts, or at once when a level has more
 * FAST candidates than its cell capacity).  Without MCV_FLAG_CHECK the call
 * only enqueues: a level that overflowed is then missing from the outputs,
 * which mcv_orb_batch_check reports. */
int mcv_orb_extract_batch_dev(mcv_ctx* ctx, int n, const uint8_t* const* d_images, const int32_t* widths,
                              const int32_t* heights, const int32_t* strides, const mcv_keypoint* const* seeds,
                              const int32_t* n_seeds, mcv_keypoint* d_kps, uint8_t* d_desc, int capacity,
                              int32_t* d_counts, int flags);
/* Waits for the last mcv_orb_extract_batch_dev on this context and returns
 * MCV_ERR_STATE when it overflowed (the pool is then grown for the next run,
 * whose outputs are complete). */
int mcv_orb_batch_check(mcv_ctx* ctx);
/* mvImagePyramid[level] of image `image` of the last mcv_orb_extract_batch_dev
 * (ORBextractor.h:72, read by Frame::ComputeStereoMatch, src/Frame.cpp:239-261),
 * left resident in HBM: a device pointer to row 0 and the row pitch in bytes.
 * blurred = 1 gives the GaussianBlur'd copy the descriptors sample instead
 * (ORBextractor.cc:874-875).  Valid until the next batched extraction of the
 * same batch shape. */
int mcv_orb_batch_pyramid_dev(mcv_ctx* ctx, int image, int level, int blurred, const uint8_t** d_level,
                              int32_t* pitch, int32_t* width, int32_t* height);

/* ---- Hamming matching --------------------------------------------------
 * Replaces Matcher::KnnMatch(const vector<Mat>&, const vector<Mat>&, int k,
 * uint n_threads) (include/Matcher.hpp:84-86, src/Matcher.cpp:245-302):
 * per query row (trainIdx0, distance0, trainIdx1, distance1), sentinel
 * (0, 999) when fewer than two trains exist.  q: nq x 32, t: nt x 32 (host). */
int mcv_match_knn2(mcv_ctx* ctx, const uint8_t* q, int nq, const uint8_t* t, int nt, int32_t* out);
/* MatchRes::FilterOrientation (src/Matcher.cpp:44-75): matches n x (query,
 * train, distance) (n <= 4096, host), angle1[query] and angle2[train] the
 * keypoint angles (degrees).  out receives the matches of the three largest
 * of 40 rotation bins, bins in the order std::sort leaves them (libstdc++'s
 * tie order), list order within a bin; *out_n their count. */
int mcv_match_filter_orientation(mcv_ctx* ctx, const int32_t* matches, int n, const float* angle1, int n1,
                                 const float* angle2, int n2, int32_t* out, int32_t* out_n);
/* MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cpp:101-150) for n
 * map points at once: the observed descriptors of map point m are rows
 * ptr[m] .. ptr[m+1] of desc (32 B each, host, at most 256 per map point);
 * best[m] = index (within the map point's rows) of the descriptor with the
 * least median distance to the others (first on ties), -1 when it has none;
 * out_desc (optional, n x 32) receives that descriptor. */
int mcv_distinctive_descriptors(mcv_ctx* ctx, const uint8_t* desc, const int32_t* ptr, int n, int32_t* best,
                                uint8_t* out_desc);
/* Launch variant of the top-2 kernel used by mcv_match_knn2* on this context
 * (0 .. 11: 0-3, 5-7 popcount kernels with scalar-loaded trains, 4 LDS-tiled,
 * 8-9 distances on the int8 matrix cores, 10-11 on the fp4 ones (train sets
 * of 65536 or more rows run variant 8); no counterpart in the reference: the
 * autotuner's choice, also settable for testing).  Every variant returns the
 * same rows. */
int mcv_match_set_variant(mcv_ctx* ctx, int variant);
int mcv_match_knn2_dev(mcv_ctx* ctx, const uint8_t* d_q, int nq, const uint8_t* d_t, int nt, int32_t* d_out);
/* Replaces MatchResKnn::FilterRatio(ratio).FilterThreshold(th)
 * (src/Matcher.cpp:100-111, :23-35): rows (queryIdx, trainIdx, distance) in
 * the reference's order (the threshold pass reorders).  knn: nq x 4 (host). */
int mcv_match_filter(mcv_ctx* ctx, const int32_t* knn, int nq, float ratio, int th, int32_t* out, int32_t* count);

/* The ratio + threshold filter of mcv_match_filter on device arrays, with
 * no allocation of any kind: no counterpart in the reference; for callers
 * that cannot allocate device memory (work under HIP graph capture, tight
 * serving loops).  d_knn: nq x 4 knn rows, d_out: nq x 3, *d_count: the
 * filtered length, all device memory; enqueued on the context's stream.
 * The filter's scratch is a fixed per-GPU array of the library (512 KiB,
 * held until the process ends) shared by every call on that GPU, so calls
 * that overlap on one GPU (other contexts or streams) race.  nq may be at
 * most 65536 (MCV_ERR_CAPACITY above). */
int mcv_match_filter_noalloc_dev(mcv_ctx* ctx, const int32_t* d_knn, int nq, float ratio, int th, int32_t* d_out,
                                 int32_t* d_count);

/* ---- rig frames: ORB on L, R, W + L-R, W-L, W-R matching ----------------
 * Frame::Frame extraction (src/Frame.cpp:118-125) followed by brute-force
 * KnnMatch -> FilterRatio -> FilterThreshold for the three camera pairs.
 * A rig object owns the plan, device outputs and the captured graph. */
/* nfeatures: NULL (the configured count for every image) or 3 * n_frames
 * per-image feature counts (frame-major L, R, W), each image then gets the
 * mnFeaturesPerLevel quotas of its own count. */
int mcv_rig_create(mcv_ctx* ctx, int n_frames, const int32_t* widths3, const int32_t* heights3,
                   const int32_t* nfeatures, float ratio, int th, mcv_rig** rig);
void mcv_rig_destroy(mcv_rig* rig);
/* Autotuner (no counterpart in the reference): times every launch variant
 * of the tunable kernels of the batched sequence on these device images
 * (median of `reps` runs), keeps the fastest for this rig and for rigs of the
 * same shape created later on this context, and re-captures the graph on the
 * next mcv_rig_run.  All variants produce identical outputs.  `report`
 * (optional) receives one line per timed variant and the choice. */
int mcv_rig_autotune(mcv_rig* rig, const uint8_t* const* d_images, int reps, char* report, int report_cap);
/* Fixes the rig's launch variants instead of tuning them: the top-2 kernel
 * (0 .. 11, as mcv_match_set_variant), the blur (bit 0: the walk, stop at the
 * level end or fixed length; bit 1: the step runs as several graphs with
 * FAST of the three largest levels and the blur on the context's side
 * stream, the blur beside the octree heap replay), the keypoint kernel (0: one keypoint per
 * wave, 1: 8, 2: 16 (the default), 3: 32) and the octree heap replay (0: top
 * of the heap in registers, 1: heap in LDS (the default)).  No counterpart in
 * the reference; every combination produces identical outputs. */
int mcv_rig_set_variants(mcv_rig* rig, int knn_variant, int blur_variant, int kp_variant, int heap_variant);
/* The rig's current launch variants (the autotuner's choice or the fixed
 * ones); any pointer may be NULL.  No counterpart in the reference. */
int mcv_rig_get_variants(const mcv_rig* rig, int32_t* knn_variant, int32_t* blur_variant, int32_t* kp_variant,
                         int32_t* heap_variant);
/* d_images: host array of 3 * n_frames device pointers (frame-major L, R, W),
 * each image dense (stride = width).  flags: MCV_FLAG_GRAPH replays the whole
 * batch as captured hipGraphs (one, or, with blur variant bit 1, several
 * joined through the context's side stream; the run's stream waits for all
 * of it); without it every image and every pair is launched on its own
 * ("plain" path). */
int mcv_rig_run(mcv_rig* rig, const uint8_t* const* d_images, int flags);
/* Stereo matching of every frame's L-R pair on the pyramids of the last
 * batched run (Frame::ComputeStereoMatch, src/Frame.cpp:150-328): row band
 * r = 10 * scale, octave +-1, disparity in [1, min(bf / baseline, 1000)),
 * ratio 0.70 / threshold 34, 11x11 SAD over 11 shifts, parabola sub-pixel
 * fit, then the median filter (1.6x / 0.4x).  Outputs per left keypoint
 * ([n_frames][capacity]): u_right and depth (-1 when unmatched) and the SAD
 * distance (-1). */
int mcv_rig_stereo(mcv_rig* rig, float bf, float baseline);
int mcv_rig_stereo_outputs(mcv_rig* rig, float** d_u_right, float** d_depth, int32_t** d_sad);
/* Waits for the last run and reports device-side overflow of the octree
 * node pool (MCV_ERR_STATE; the pool is grown for the next run). */
int mcv_rig_check(mcv_rig* rig);
/* The batched sequence launched eagerly with an event after every launch:
 * ms[i] = duration of launch i (ms), names[i] = its kernel (static strings),
 * n = number of launches.  For kernel timing only. */
int mcv_rig_run_timed(mcv_rig* rig, const uint8_t* const* d_images, float* ms, const char** names, int cap,
                      int32_t* n);
/* mvImagePyramid[level] of image `image` (frame-major L, R, W) of the last
 * batched mcv_rig_run, resident in HBM (as mcv_orb_batch_pyramid_dev). */
int mcv_rig_pyramid_dev(mcv_rig* rig, int image, int level, int blurred, const uint8_t** d_level, int32_t* pitch,
                        int32_t* width, int32_t* height);
/* Octree sizes of the last run (summed over every level of every image):
 * FAST candidates entering DistributeOctTree, quadtree node-table entries
 * built, nodes emitted.  No counterpart in the reference; for measurement. */
int mcv_rig_octree_stats(mcv_rig* rig, int64_t* candidates, int64_t* table_entries, int64_t* nodes);
/* Device outputs: keypoints [3n][capacity], descriptors [3n][capacity][32],
 * counts [3n], matches [3n pairs][capacity][3] (query, train, distance),
 * match counts [3n]; knn rows [3n pairs][capacity][4]. */
int mcv_rig_outputs(mcv_rig* rig, mcv_keypoint** d_kps, uint8_t** d_desc, int32_t** d_counts, int32_t** d_matches,
                    int32_t** d_match_counts, int32_t** d_knn, int32_t* capacity);

/* ---- projection-guided window search -----------------------------------
 * Object::AssignFeaturesToGrid + ProjectBunchMapPoints + GetFeaturesInArea +
 * KnnMatch/FilterRatio/FilterThreshold for every map point in order
 * (src/Object.cpp:182-236, :249-308).  Host arrays.  Per map point the
 * output row is (kp0, d0, kp1, d1, n_candidates, match) with keypoint
 * indices (-1 for the 999 sentinel) and match = accepted keypoint or -1.
 * R, t: world-to-camera (row-major 3x3 and 3, f32); cam: fx, fy, cx, cy;
 * scales: per-level scale factors (nlevels entries); every plevel must lie in
 * [0, nlevels) (MCV_ERR_INVALID otherwise).  ProjectBunchMapPoints applies
 * no level filter, so the keypoints' octaves are not an input. */
int mcv_window_search(mcv_ctx* ctx, const float* points, const uint8_t* pdesc, const int32_t* plevel, int n_points,
                      const float* R, const float* t, const float* cam, const float* scales, int nlevels, float radius,
                      float ratio, int th, const float* kxy, const uint8_t* kdesc, int n_kps, int width, int height,
                      int32_t* out);
/* Device variant over several cameras at once: arrays are device pointers,
 * cameras index blocks of keypoints; out is [n_cams][n_points][6].  A map
 * point whose level is outside [0, nlevels) gets the row (-1, 999, -1, 999,
 * -1, -1) (n_candidates = -1 marks it). */
int mcv_window_search_dev(mcv_ctx* ctx, int n_cams, const float* d_points, const uint8_t* d_pdesc,
                          const int32_t* d_plevel, int n_points, const float* Rt /* host: n_cams x 12 */,
                          const float* cams /* host: n_cams x 4 */, const int32_t* sizes /* host: n_cams x 2 */,
                          const float* scales, int nlevels, float radius, float ratio, int th,
                          const float* const* d_kxy, const uint8_t* const* d_kdesc,
                          const int32_t* n_kps /* host */, int32_t* d_out);
/* Window query by position, the loop of Tracker::Wnd_Track
 * (src/Tracker.cpp:339-353): for each of n query positions d_uv (n x 2) with
 * descriptor d_qdesc (n x 32), GetFeaturesInArea(x, y, radius, min_level,
 * max_level) (src/Object.cpp:263-308, level filter on d_koct as there: active
 * when min_level > 0 or max_level >= 0) over the keypoints d_kxy / d_koct /
 * d_kdesc of a width x height image, then KnnMatch -> FilterRatio(ratio) ->
 * FilterThreshold(th) of the query against the candidates.  Device arrays;
 * d_out is n x 6 as mcv_window_search.  (FilterOrientation of one match keeps
 * it, so Wnd_Track's accepted match is the row's `match`.) */
int mcv_window_search_uv_dev(mcv_ctx* ctx, const float* d_uv, const uint8_t* d_qdesc, int n, float radius,
                             int min_level, int max_level, float ratio, int th, const float* d_kxy,
                             const int32_t* d_koct, const uint8_t* d_kdesc, int n_kps, int width, int height,
                             int32_t* d_out);

/* ---- bag of words (DBoW3, vendored by the reference: modules/DBow3) -----
 * A vocabulary tree on the device: node 0 the root, children of node i
 * child_ids[child_ptr[i] .. child_ptr[i + 1]) in DBoW3's children order, n x
 * 32 B node descriptors, word_id (-1 for inner nodes) and weight per node,
 * L the vocabulary's depth parameter (m_L), weighting 0 TF_IDF, 1 TF, 2 IDF,
 * 3 BINARY, scoring 0 L1, 1 L2, 2 chi-square, 3 KL, 4 Bhattacharyya, 5 dot
 * product (DBoW3.h's enums).  Replaces Vocabulary's loaded tree
 * (Vocabulary.h:56-60 m_nodes, m_words). */
typedef struct mcv_bow_vocab mcv_bow_vocab;
int mcv_bow_vocab_create(mcv_ctx* ctx, int n_nodes, const int32_t* child_ptr, const int32_t* child_ids,
                         const uint8_t* node_desc, const int32_t* word_id, const double* weight, int L,
                         int weighting, int scoring, mcv_bow_vocab** vocab);
void mcv_bow_vocab_destroy(mcv_bow_vocab* vocab);
/* Vocabulary::transform(feature, word_id, weight, nid, levelsup)
 * (modules/DBow3/src/Vocabulary.cpp:641-672) for n device descriptors
 * (n x 32 B): the leaf reached by descending to the child at the least
 * Hamming distance (the first of equal distances), its word id and weight,
 * and the node passed at level L - levelsup (0, the root, when that level is
 * <= 0 or below the leaf).  Device outputs of n entries. */
int mcv_bow_transform_dev(mcv_bow_vocab* vocab, const uint8_t* d_feats, int n, int levelsup, int32_t* d_word,
                          double* d_weight, int32_t* d_nid);
/* Vocabulary::transform(features, BowVector, FeatureVector, levelsup)
 * (Vocabulary.cpp:572-629, called by Object::ComputeBow, src/Object.cpp:
 * 238-247, with levelsup 4) for n host descriptors: the per-feature outputs
 * (any may be NULL), the BowVector as ascending word ids and values
 * (*n_words of them; capacity n), and the FeatureVector as ascending node ids,
 * CSR offsets (*n_nodes + 1) and feature indices (capacity n each). */
int mcv_bow_transform(mcv_bow_vocab* vocab, const uint8_t* feats, int n, int levelsup, int32_t* word, double* weight,
                      int32_t* nid, int32_t* bv_words, double* bv_vals, int32_t* n_words, int32_t* fv_nodes,
                      int32_t* fv_ptr, int32_t* fv_feats, int32_t* n_nodes);
/* Matcher::DBowMatch (src/Matcher.cpp:146-193): descriptors d1 (n1 x 32) and
 * d2 (n2 x 32) with their FeatureVectors (m nodes ascending, CSR offsets,
 * feature indices), host arrays.  out receives a row (query, train0, d0,
 * train1, d1) per feature of image 1 in a node both hold where image 2 has
 * at least two features, nodes ascending and features in list order (the
 * reference's MatchResKnn order; capacity = the features of image 1);
 * *count the rows. */
int mcv_bow_match(mcv_ctx* ctx, const uint8_t* d1, int n1, int m1, const int32_t* nodes1, const int32_t* ptr1,
                  const int32_t* feats1, const uint8_t* d2, int n2, int m2, const int32_t* nodes2,
                  const int32_t* ptr2, const int32_t* feats2, int32_t* out, int32_t* count);

/* ---- bundle-adjustment linearization -----------------------------------
 * The batched seam of BlockSolver::buildSystem / SparseOptimizer::
 * computeActiveErrors (ThirdParty/g2o/g2o/core/block_solver.hpp:463-520,
 * core/sparse_optimizer.cpp:63-90) for EdgeSE3XYZ2Camera edges
 * (include/OptimizeTypes.hpp:70-117) added by BAoptimizer::addEdge /
 * addEdgeToBody (src/BAoptimizer.cpp:45-65, :90-108). */
int mcv_ba_create(mcv_ctx* ctx, int n_obs, int n_points, int n_frames, int n_cams, mcv_ba** ba);
void mcv_ba_destroy(mcv_ba* ba);
/* Observations (host): uv (M x 2 f32), point, frame, camera indices (M),
 * inverse level sigma^2 (M f32: mvInvLevelSigma2[level]). */
int mcv_ba_set_observations(mcv_ba* ba, const float* uv, const int32_t* point, const int32_t* frame,
                            const int32_t* cam, const float* inv_sigma2);
/* State (host, FP64 as g2o keeps it): points P x 3; frames F x 7 body-from-
 * world (qx qy qz qw tx ty tz); cams C x 12 (fx fy cx cy, camera-from-body
 * qx qy qz qw tx ty tz, Huber delta). */
int mcv_ba_set_state(mcv_ba* ba, const double* points, const double* frames, const double* cams);
/* Linearize every edge into 22 device planes of M floats: e_u, e_v, chi2,
 * Huber weight, J_pose (2x6 row-major), J_point (2x3 row-major).  mode:
 * MCV_BA_FULL every plane; MCV_BA_RESIDUAL the first four only (the LM
 * trials' SparseOptimizer::computeActiveErrors, core/sparse_optimizer.cpp:
 * 63-90); MCV_BA_POSE_ONLY all but J_point: the edge with a fixed point,
 * whose residual and J_pose are those of EdgeSE3XYZ2CameraOnlyPose
 * (include/OptimizeTypes.hpp:20-68; PoseEstimation::PoseOptimization fixes
 * every point, src/PoseEstimation.cpp:45-77). */
#define MCV_BA_FULL 0
#define MCV_BA_RESIDUAL 1
#define MCV_BA_POSE_ONLY 2
int mcv_ba_linearize(mcv_ba* ba, int mode);
/* Launch shape of mcv_ba_linearize (0 .. 7: workgroup size, grid cap,
 * observations in flight per thread, streaming stores; no counterpart in the
 * reference).  Every variant computes the same values per observation. */
int mcv_ba_set_variant(mcv_ba* ba, int variant);
/* Times every launch shape of mcv_ba_linearize on this problem (median of
 * `reps` launches after one warm-up) and keeps the fastest; no counterpart in
 * the reference.  Leaves the planes linearized.  `report` (optional) receives
 * one line per variant and the choice. */
int mcv_ba_autotune(mcv_ba* ba, int reps, char* report, int report_cap);
/* Same outputs with the plain launch pattern: one launch per (frame, camera)
 * edge group, the granularity at which g2o's per-edge loop
 * (core/block_solver.hpp:488-494) would be replaced group by group.  For the
 * batched-vs-plain comparison. */
int mcv_ba_linearize_plain(mcv_ba* ba, int residual_only);
/* Stereo edges, EdgeStereoSE3ProjectXYZ (include/OptimizeTypes.hpp:119-157,
 * src/OptimizeTypes.cpp:12-58; added by BAoptimizer::addEdgeStereo,
 * src/BAoptimizer.cpp:67-87): obs Ms x 3 f32 (u, v, u_right), point and frame
 * indices, inverse level sigma^2; cam = {fx, fy, cx, cy, bf, Huber delta}.
 * The pose vertex is the frame state of mcv_ba_set_state (left camera = body). */
int mcv_ba_set_stereo(mcv_ba* ba, int Ms, const float* obs, const int32_t* point, const int32_t* frame,
                      const float* inv_sigma2, const double* cam);
/* Linearize the stereo edges into 32 device planes of Ms floats: e_u, e_v,
 * e_ur, chi2, Huber weight, J_pose (3x6 row-major), J_point (3x3).  mode as
 * mcv_ba_linearize (MCV_BA_RESIDUAL: the first five planes). */
int mcv_ba_linearize_stereo(mcv_ba* ba, int mode);
/* Copy stereo planes [first, first + count) to host (count x Ms floats). */
int mcv_ba_read_stereo_planes(mcv_ba* ba, int first, int count, float* host);
/* Outlier test of PoseEstimation::FilterCallBack_Chi2 (src/PoseEstimation.cpp:
 * 132-151) on the last linearization: flag = chi2 >= threshold (the reference
 * uses 5.991 for edges of the left camera object, 7.815 for the right).
 * Host outputs of M (mono) and Ms (stereo) bytes; either may be null. */
int mcv_ba_outliers(mcv_ba* ba, float thr_mono, float thr_stereo, uint8_t* flags_mono, uint8_t* flags_stereo);
/* Version of the planes: changes with every mcv_ba_set_state,
 * mcv_ba_set_observations and linearization.  A caller that keeps the
 * linearization for later (the autograd backward) records it and checks that
 * it is unchanged before mcv_ba_adjoint_dev.  No counterpart in the reference. */
int mcv_ba_generation(mcv_ba* ba, int64_t* gen);
/* Device pointer of the planes (plane k at d_planes + k * stride). */
int mcv_ba_planes(mcv_ba* ba, float** d_planes, int64_t* stride);
/* Copy planes [first, first + count) to host (count x M floats). */
int mcv_ba_read_planes(mcv_ba* ba, int first, int count, float* host);
/* Device-to-device copy of planes [first, first + count) (enqueued). */
int mcv_ba_copy_planes_dev(mcv_ba* ba, int first, int count, float* d_dst);
/* Adjoint of the residual: grad_pose[F x 6] = sum J_pose^T g and
 * grad_point[P x 3] = sum J_point^T g over the last linearization, with
 * d_g an M x 2 device array; outputs are device arrays. */
int mcv_ba_adjoint_dev(mcv_ba* ba, const float* d_g, float* d_grad_pose, float* d_grad_point);

/* ---- Schur complement of the point blocks ------------------------------
 * BlockSolver::buildSystem + the Schur part of BlockSolver::solve with the LM
 * damping of setLambda (ThirdParty/g2o/g2o/core/block_solver.hpp:315-449,
 * 463-520): from the last linearization of the active mono and stereo edges
 * (constructQuadraticForm, core/base_binary_edge.hpp:64-136), in FP32 with
 * fixed summation orders, the reduced camera system S (n x n, n = 6 x free
 * frames, frames numbered in order) and b_schur.  Fixed frames
 * (VertexSE3Expmap::setFixed, src/BAoptimizer.cpp:31) get no pose block;
 * fixed points (BAoptimizer::addMapppint(mp, fixed), src/BAoptimizer.cpp:
 * 34-43) are not eliminated and their edges enter H_pp and b_p only. */
int mcv_ba_set_fixed(mcv_ba* ba, const int32_t* frame_fixed);
/* point_fixed: P flags, or NULL for none. */
int mcv_ba_set_point_fixed(mcv_ba* ba, const int32_t* point_fixed);
/* Edge levels of g2o (OptimizableGraph::Edge::setLevel; an edge at level 1
 * takes no part in the optimization, src/PoseEstimation.cpp:104-118):
 * active_mono (M bytes) / active_stereo (Ms bytes), NULL = all active.  They
 * select the edges of mcv_ba_schur, mcv_ba_robust_chi2 and mcv_ba_optimize. */
int mcv_ba_set_active(mcv_ba* ba, const uint8_t* active_mono, const uint8_t* active_stereo);
/* enabled = 0 drops the Huber kernels of every edge (setRobustKernel(0),
 * src/PoseEstimation.cpp:117): weight 1 and rho(chi2) = chi2. */
int mcv_ba_set_robust(mcv_ba* ba, int enabled);
int mcv_ba_schur(mcv_ba* ba, float lambda);
/* Host copy of S (row-major, full symmetric) and b_schur as double. */
int mcv_ba_read_schur(mcv_ba* ba, double* S, double* b_schur, int32_t* n);
/* FP64 dense Cholesky solve S x = b on the host (stands in for g2o's
 * LinearSolverEigen, solvers/eigen/linear_solver_eigen.h:55). */
int mcv_solve_llt(int n, const double* S, const double* b, double* x);

/* ---- Levenberg-Marquardt ------------------------------------------------
 * SparseOptimizer::activeRobustChi2 (core/sparse_optimizer.cpp:102): the
 * sum over the active edges of the last linearization (mono, and stereo when
 * set) of the Huber rho(chi2) (core/robust_kernel_impl.cpp:65-78), reduced
 * on the GPU in FP64 in a fixed order. */
int mcv_ba_robust_chi2(mcv_ba* ba, double* chi2);
/* SparseOptimizer::optimize(iterations) with OptimizationAlgorithmLevenberg
 * (core/optimization_algorithm_levenberg.cpp:56-159; BAoptimizer::optimize,
 * src/BAoptimizer.cpp:110-114) on the state of mcv_ba_set_state: per
 * iteration the GPU linearizes every active edge and sums rho(chi2), builds
 * the Schur complement with lambda (initially tau times the largest diagonal
 * entry of H; g2o's tau is 1e-5 and its trial limit 10), the host solves the
 * reduced system in FP64, the GPU back-substitutes the points and applies
 * the update (points += x_l on the device, keyframes exp(x_p) * T on the
 * host), a residual-only linearization gives the trial's chi2, and lambda
 * follows g2o's rules (step accepted when rho > 0: lambda *= max(1/3,
 * min(2/3, 1 - (2 rho - 1)^3)); rejected: the state is restored, lambda *=
 * ni, ni *= 2).  stats (optional, iterations x 4): chi2 before the
 * iteration, after it, lambda, trials.  *done: iterations run (a Terminate
 * ends the run early, as in g2o).  The planes afterwards hold the residuals
 * (e, chi2, w) of the final state, for mcv_ba_outliers. */
int mcv_ba_optimize(mcv_ba* ba, int iterations, double tau, int max_trials, double* stats, int32_t* done);
/* The estimates (host, FP64): points P x 3 and frames F x 7 as
 * mcv_ba_set_state takes them (BAoptimizer::eval / recovery_all,
 * src/BAoptimizer.cpp:116-145); either may be NULL. */
int mcv_ba_get_state(mcv_ba* ba, double* points, double* frames);

#ifdef __cplusplus
}
#endif
#endif /* MCVSLAM_GPU_H */
