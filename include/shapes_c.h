/*
 * shapes_c.h - C ABI of the MI355X (gfx950) normal-estimation pipeline.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns an
 * int status (0 = SX_OK) and never throws.  sx_last_error() returns a
 * thread-local message for the last failing call on the calling thread.
 *
 * Conventions
 *   - Pointers named d_* or documented as "device" are device (HBM) pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Compute entry points are asynchronous on that stream and never
 *     synchronise, unless documented otherwise.
 *   - Stage entry points that need scratch memory take (void* ws, size_t*
 *     ws_bytes): call once with ws == NULL to receive the size, then again
 *     with a device buffer of at least that size (the two-phase pattern cub
 *     uses at src/RadixTree.cu:318-331 of the reference).
 *   - Data-dependent counts (unique points, octree nodes) stay on the device
 *     (int32 at a caller-visible address) so that a chain of calls never waits
 *     for the host; kernels size their grids from a host capacity and read the
 *     true count from device memory.
 *
 * Reference files cited below are those of Wehrdo/ShapeExtraction.
 */
#ifndef SHAPES_C_H_
#define SHAPES_C_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SX_ABI_VERSION 2

/* Status codes. */
#define SX_OK 0
#define SX_EINVAL 1   /* invalid argument (null pointer, bad size, unsupported k) */
#define SX_ENOMEM 2   /* device allocation failed or workspace too small          */
#define SX_EHIP 3     /* HIP runtime error                                          */
#define SX_ERANGE 4   /* a data-dependent size exceeded the capacity given          */
#define SX_EDATA 5    /* invalid input data (non-finite coordinates)                */
#define SX_ESTATE 6   /* object used before it was built                            */
#define SX_EINTERNAL 7 /* a device-side invariant broke (a sort look-back gave up);
                          the results of that build are invalid                    */

/* Karras root rule.  The reference gives the root l = N-1 with N = n_unique-1
 * (src/RadixTree.cu:128-131 with :353-362), so its root spans leaves
 * [0, n_unique-2].  TEXTBOOK spans [0, n_unique-1].  The two agree except when
 * the last code alone forms the root's right subtree or that subtree is exactly
 * the last two leaves. */
#define SX_ROOT_TEXTBOOK 0
#define SX_ROOT_REFERENCE 1

/* kNN modes. */
#define SX_KNN_PLAIN 0    /* exact nearest-first descent from the root            */
#define SX_KNN_RADIUS 1   /* exact descent seeded with a per-query radius          */

/* Index algorithm choices (results are identical; speed differs). */
#define SX_KNN_ALGO_TILES 0     /* wave-cooperative tiles + per-lane queue (default) */
#define SX_KNN_ALGO_LANE 1      /* one lane per query, stackless descent            */
#define SX_SORT_ONESWEEP 0      /* single-pass-per-digit radix sort                 */
#define SX_SORT_THREE_KERNEL 1  /* histogram, scan, scatter per digit               */
#define SX_SORT_HYBRID 2        /* top digits by onesweep, then small buckets sorted
                                   in LDS, large ones by onesweep (default)         */

/* Normal precision modes. */
#define SX_PREC_FP32 0    /* compensated FP32 covariance + closed-form FP32 eigen,
                             FP64 Jacobi for near-repeated smallest eigenvalues   */
#define SX_PREC_FP64 1    /* FP64 covariance + FP64 Jacobi for every query         */

/* Octree node; byte-identical to OT::OTNode (src/Octree.hpp:19-35): 56 bytes. */
typedef struct sx_otnode {
    int32_t children[8];
    float corner[3];
    float cell_size;
    int32_t child_node_mask;
    int32_t child_leaf_mask;
} sx_otnode;

/* ------------------------------------------------------------------------ */
/* Library and device runtime (plumbing for harnesses in any language).      */
/* ------------------------------------------------------------------------ */
const char* sx_last_error(void);
int sx_abi_version(void);
int sx_device_count(int* n);
int sx_set_device(int device);
int sx_device_name(int device, char* buf, size_t buf_len);
int sx_malloc(void** d_ptr, size_t bytes);
int sx_free(void* d_ptr);
int sx_host_malloc(void** h_ptr, size_t bytes);   /* pinned host memory */
int sx_host_free(void* h_ptr);
int sx_memcpy_htod(void* d_dst, const void* h_src, size_t bytes, void* stream);
int sx_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes, void* stream);
int sx_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes, void* stream);
int sx_fill_u8(void* d_dst, uint8_t value, size_t bytes, void* stream);  /* kernel fill, graph-safe */
int sx_stream_create(void** stream);
int sx_stream_destroy(void* stream);
int sx_stream_sync(void* stream);
int sx_device_sync(void);
int sx_event_create(void** event);
int sx_event_destroy(void* event);
int sx_event_record(void* event, void* stream);
int sx_event_elapsed_ms(void* start, void* stop, float* ms);

/* ------------------------------------------------------------------------ */
/* Synthetic workloads (host code; deterministic counter-based RNG).          */
/* ------------------------------------------------------------------------ */
uint64_t sx_fnv64(const void* data, size_t bytes);
int sx_gen_sphere(uint64_t seed, int64_t n, float noise_sigma, float* x, float* y, float* z);
int sx_gen_scene(uint64_t seed, int64_t n, int n_prims, float box, float noise_sigma, float outlier_frac,
                 float* x, float* y, float* z);
int sx_gen_gmm_scene(uint64_t seed, int64_t n, int n_prims, int n_components, float box, float noise_sigma,
                     float max_density_ratio, float* x, float* y, float* z, float* base_fraction_out);
int sx_gen_batch_sizes(uint64_t seed, int n_clouds, int64_t lo, int64_t hi, int64_t* sizes);
int sx_gen_degenerate(uint64_t seed, int64_t n, int n_per_class_structs, float offset, float box,
                      float* x, float* y, float* z, uint8_t* cls, float* dir);

/* ------------------------------------------------------------------------ */
/* Stage entry points (one per reference kernel / library call).              */
/* ------------------------------------------------------------------------ */

/* Cube bounds of the cloud: d_bounds = {min over x,y,z; max over x,y,z;
 * max-min}.  Replaces the six cub DeviceReduce::Min/Max calls and the host
 * min/max of RadixTree::encodePoints (src/RadixTree.cu:208-241).  Non-finite
 * coordinates set *d_bad (device int32, may be NULL) to their count. */
int sx_bbox(const float* d_x, const float* d_y, const float* d_z, int64_t n, float* d_bounds, int32_t* d_bad,
            void* ws, size_t* ws_bytes, void* stream);

/* Morton codes of the points under the cube d_bounds; code_bits is 63 (the
 * reference's CODE_LEN, src/MortonUtils.hpp:10) or 30 (10 bits per axis).
 * Replaces pointToCode + makeCodes (src/MortonUtils.cu:7-19,
 * src/RadixTree.cu:21-39).  A zero-extent cube gives code 0. */
int sx_morton_encode(const float* d_x, const float* d_y, const float* d_z, int64_t n, const float* d_bounds,
                     int code_bits, uint64_t* d_codes, void* stream);

/* Decoded points (12-byte AoS x,y,z) of codes[0..count), count = *d_count if
 * d_count != NULL else n_max.  Replaces codeToPoint + decodePoints
 * (src/MortonUtils.cu:21-28, src/Octree.cu:14-24). */
int sx_morton_decode(const uint64_t* d_codes, const int32_t* d_count, int64_t n_max, const float* d_bounds,
                     int code_bits, float* d_points, void* stream);

/* Stable ascending LSD radix sort of (key, value) pairs over key bits
 * [0, key_bits).  d_vals_in == NULL sorts the values 0..n-1 (the sort
 * permutation).  Replaces cub DeviceRadixSort::SortKeys
 * (src/RadixTree.cu:316-335), which sorts keys only. */
int sx_sort_pairs(const uint64_t* d_keys_in, const uint32_t* d_vals_in, int64_t n, int key_bits,
                  uint64_t* d_keys_out, uint32_t* d_vals_out, void* ws, size_t* ws_bytes, void* stream);
/* Status of the last sx_sort_pairs that used workspace `ws` for n keys
 * (synchronises `stream`): SX_EINTERNAL when one of its single-pass tiles
 * gave up waiting for an earlier tile's prefix (the forward-progress
 * invariant broke and the output is invalid), else SX_OK.  No reference
 * counterpart (cub reports nothing either, src/RadixTree.cu:316-335). */
int sx_sort_check(const void* ws, int64_t n, void* stream);
/* Spins a sort tile waits for an earlier tile before giving up (default
 * 2^22).  A process-wide setting read at launch; a test hook for the
 * give-up path above (0 gives up at the first wait, and every kernel still
 * exits). */
int sx_sort_set_spin_limit(uint32_t spins);

/* Duplicate removal over sorted codes: d_unique[0..U) are the distinct codes,
 * d_rep[u] the value (input index) of the first element of run u,
 * d_unique_of_sorted[i] the run of sorted element i (nullable), *d_count = U.
 * Replaces findDups + InclusiveSum + moveDups (src/RadixTree.cu:71-97,258-303). */
int sx_unique(const uint64_t* d_sorted, const uint32_t* d_vals, int64_t n, uint64_t* d_unique, uint32_t* d_rep,
              int32_t* d_unique_of_sorted, int32_t* d_count, void* ws, size_t* ws_bytes, void* stream);

/* Karras binary radix tree over the U = *d_count sorted unique codes (arrays
 * sized >= U, capacity n_max).  Output arrays follow RT::Nodes
 * (src/RadixTree.hpp:17-34): prefix_n, left_child, has_leaf_left/right,
 * parent (parent[0] = -1).  Replaces constructTree (src/RadixTree.cu:99-194).
 * d_node_end (nullable) receives the far end j of each internal node's leaf
 * range [min(i,j), max(i,j)]; d_leaf_parent (nullable) receives for every
 * leaf u the internal node holding it, times 2, plus 1 for a right child. */
int sx_build_radix_tree(const uint64_t* d_codes, const int32_t* d_count, int64_t n_max, int code_bits,
                        int root_mode, uint8_t* d_prefix_n, int32_t* d_left_child, uint8_t* d_has_leaf_left,
                        uint8_t* d_has_leaf_right, int32_t* d_parent, int32_t* d_node_end,
                        int32_t* d_leaf_parent, void* stream);

/* Octree from the radix tree: edge counts, own scan, node chains and leaf
 * links, with OT::OTNode layout and node numbering identical to the
 * reference (src/Octree.cu:38-50,65-161,213-261).  oct_capacity bounds the
 * node array; *d_oct_count receives the node count (SX_ERANGE is reported by
 * sx_check_counts when it exceeds the capacity; nodes beyond it are not
 * written). */
int sx_build_octree(const uint64_t* d_codes, const int32_t* d_count, int64_t n_max, int code_bits,
                    const float* d_bounds, const uint8_t* d_prefix_n, const int32_t* d_left_child,
                    const uint8_t* d_has_leaf_left, const uint8_t* d_has_leaf_right, const int32_t* d_parent,
                    sx_otnode* d_nodes, int64_t oct_capacity, int32_t* d_oct_count, void* ws, size_t* ws_bytes,
                    void* stream);

/* Reads device counts (synchronises `stream`) and validates them against
 * capacities; any pointer may be NULL. */
int sx_check_counts(const int32_t* d_count, int64_t count_capacity, const int32_t* d_oct_count,
                    int64_t oct_capacity, const int32_t* d_bad, int64_t* count_out, int64_t* oct_count_out,
                    void* stream);

/* ------------------------------------------------------------------------ */
/* Index object: one cloud's full spatial index, built without host syncs.   */
/* Replaces the RT::RadixTree and OT::Octree constructors                    */
/* (src/RadixTree.cu:305-382, src/Octree.cu:213-319).                        */
/* ------------------------------------------------------------------------ */
typedef struct sx_index sx_index;

typedef struct sx_index_view {
    int64_t n_input;     /* points given to the last build                  */
    int64_t n_unique;    /* RadixTree::n_pts after duplicate removal        */
    int64_t n_oct;       /* Octree::n_nodes                                 */
    float min_coord;     /* RadixTree::min_coord                            */
    float max_coord;     /* RadixTree::max_coord                            */
    int code_bits;
    /* device arrays */
    const uint64_t* codes;              /* [n_unique] RT::Nodes::mortonCode          */
    const uint8_t* has_leaf_left;       /* [n_unique]                                */
    const uint8_t* has_leaf_right;      /* [n_unique]                                */
    const uint8_t* prefix_n;            /* [n_unique]                                */
    const int32_t* left_child;          /* [n_unique]                                */
    const int32_t* parent;              /* [n_unique]                                */
    const sx_otnode* oct_nodes;         /* [n_oct]    Octree::u_nodes                */
    const float* points;                /* [3*n_unique] Octree::u_points (12-B AoS)  */
    const uint32_t* sorted_perm;        /* [n_input] stable sort permutation         */
    const uint32_t* rep;                /* [n_unique] first input index of each code */
    const int32_t* unique_of_input;     /* [n_input] unique id of each input point   */
    const int32_t* d_counts;            /* device {n_unique, n_oct, n_bad}           */
    const float* d_bounds;              /* device {min, max, range}                  */
} sx_index_view;

/* capacity: maximum input points; oct_factor: octree node capacity as a
 * multiple of capacity (<= 0 selects 3.0). */
int sx_index_create(sx_index** out, int64_t capacity, int code_bits, int root_mode, float oct_factor);
int sx_index_destroy(sx_index* index);
/* Builds the index from device SoA coordinates; asynchronous on `stream`. */
int sx_index_build(sx_index* index, const float* d_x, const float* d_y, const float* d_z, int64_t n,
                   void* stream);
/* Builds the index from n strictly increasing Morton codes (device, at most
 * the index's code bits wide) of points under the cube [min_coord,
 * min_coord + range]: no bounds, encode, sort or duplicate removal; the
 * permutation and representatives are the identity, the points are the
 * decoded codes.  The path of the reference's Octree host-data constructor
 * (src/Octree.cu:179-192), whose points are the decoded unique points of an
 * earlier build in Morton order.  Codes that break the order are reported by
 * sx_index_get_view (SX_EDATA) and leave the index empty. */
int sx_index_build_codes(sx_index* index, const uint64_t* d_codes, int64_t n, float min_coord, float range,
                         void* stream);
/* Orders `stream` after the last build (whatever stream it ran on),
 * synchronises it, validates counts and fills *view.  SX_EINTERNAL when the
 * build's sort failed (see sx_sort_check). */
int sx_index_get_view(sx_index* index, sx_index_view* view, void* stream);
/* Rows in input order: output row i (i < n_input) = row unique_of_input[i] of
 * d_rows, rows of row_bytes (a multiple of 4) per unique point as written by
 * sx_normals / sx_knn for the whole index.  Input points that share a code
 * share its row.  The reference returns unique points in Morton order only
 * (src/main.cu:119-120); this is the scatter back to the caller's order. */
int sx_index_to_input_order(sx_index* index, const void* d_rows, int64_t row_bytes, void* d_out, void* stream);
/* Bucket size of the kNN descent: a child holding at most `bucket` points is
 * scanned linearly instead of descended (default 32; tuning knob, no effect
 * on results).  No reference counterpart (the reference descends to single
 * points, src/Octree.hpp:160-178). */
int sx_index_set_bucket(sx_index* index, int bucket);
/* Algorithms used by this index (SX_KNN_ALGO_*, SX_SORT_*). */
int sx_index_set_algorithms(sx_index* index, int knn_algo, int sort_algo);
/* Number of queries the kNN could not settle on its fast path (tile search,
 * or the seeded search of SX_KNN_RADIUS) and searched again one lane per
 * query, since the last reset (synchronises `stream`). */
int sx_index_knn_reruns(sx_index* index, int64_t* reruns, int reset, void* stream);

/* Query entry points (sx_knn, sx_knn_points, sx_normals, sx_estimate_normals,
 * sx_radius_*) order their stream after the last build.  Query positions are
 * checked against the unique count once a view has read it, else against the
 * input count; rows of positions at or beyond the unique count are written as
 * -1 (neighbours) or NaN (normals).  An index is used from one stream at a
 * time: its calls share the deferred-query queue and counters, so calls that
 * overlap on different streams race. */

/* Work statistics of the tile kNN (SX_KNN_ALGO_TILES): while enabled, every
 * sx_knn call adds the candidate points its tiles broadcast to their 64
 * queries (the kernel's dominant work) to a device counter;
 * sx_index_knn_candidates reads it (synchronises `stream`).  Disabled by
 * default (one atomic per tile).  No reference counterpart. */
int sx_index_set_knn_stats(sx_index* index, int enable);
int sx_index_knn_candidates(sx_index* index, int64_t* candidates, int reset, void* stream);

/* Exact k nearest neighbours of the index points [q_start, q_start+n_q)
 * among all index points, k in {8, 16, 24, 32}.  Row q of d_out holds k
 * indices ascending by (squared distance, index); the query itself is
 * included; rows are padded with their first entry when fewer than k points
 * exist (src/Octree.hpp:181-191).  mode SX_KNN_RADIUS seeds each query with
 * the squared radius d_radius2[q] (device, may under- or over-estimate; the
 * result stays exact).  Replaces Octree::deviceKnnSearch<k> and
 * knnSearchKernel<k> (src/Octree.hpp:136-212). */
int sx_knn(sx_index* index, int64_t q_start, int64_t n_q, int k, int mode, const float* d_radius2,
           int32_t* d_out, void* stream);

/* Exact k nearest index points of n_q arbitrary query points (12-B AoS xyz,
 * device); rows as in sx_knn.  The query API the reference's host
 * Octree::knnSearch<k> aimed at (src/Octree.hpp:214-242, whose call passes
 * its arguments in the wrong order at :232). */
int sx_knn_points(sx_index* index, const float* d_queries, int64_t n_q, int k, int32_t* d_out, void* stream);

/* Normals of the index points [q_start, q_start+n_q) from their k-neighbour
 * rows d_nn (as written by sx_knn): eigenvector of the smallest covariance
 * eigenvalue, unit length, flipped so that (viewpoint - p) . n >= 0.  Output
 * is 12-B AoS.  Replaces findCovariancesKernel + cusolverDnSsyevjBatched +
 * normalsFromEigenVectors (src/NormalEstimation.hpp:96-192,235-315).
 * d_flags (nullable, uint8 per query) receives 1 where the FP64 path ran. */
int sx_normals(sx_index* index, const int32_t* d_nn, int64_t q_start, int64_t n_q, int k, int precision,
               float vx, float vy, float vz, float* d_normals, uint8_t* d_flags, void* stream);

/* kNN + normals for [q_start, q_start+n_q), viewpoint (0,0,0): the body of
 * NormalEstimation::estimateNormals<k>(octree, start, n)
 * (src/NormalEstimation.hpp:196-335).  d_nn may be NULL (then an internal
 * buffer holds the neighbour rows). */
int sx_estimate_normals(sx_index* index, int64_t q_start, int64_t n_q, int k, int knn_mode,
                        const float* d_radius2, int precision, int32_t* d_nn, float* d_normals, void* stream);

/* sx_estimate_normals without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The neighbour rows
 * live in a fixed 64 MiB device array of the library (one per GPU, held until
 * the process ends), so n_q * k <= 2^24 (e.g. 1,048,576 queries at k = 16);
 * larger ranges are split by the caller (q_start, n_q).  Every call on a GPU
 * shares that array: calls that overlap on one GPU (different streams or
 * threads) race.  Results equal sx_estimate_normals with d_nn == NULL. */
int sx_estimate_normals_noalloc(sx_index* index, int64_t q_start, int64_t n_q, int k, int knn_mode,
                                const float* d_radius2, int precision, float* d_normals, void* stream);

/* ------------------------------------------------------------------------ */
/* Learned search radius (no reference counterpart: the reference search has  */
/* no radius, src/Octree.hpp:152-155).  A 8-64-64-1 MLP, BF16 weights, FP32   */
/* accumulation on the matrix cores; trained on exact k-th neighbour          */
/* distances from sx_knn.  Parameters are a flat FP32 vector: W1[64][8], b1,  */
/* W2[64][64] (out, in), b2, w3[64], b3 (sx_mlp_num_params() floats).         */
/* ------------------------------------------------------------------------ */
typedef struct sx_mlp sx_mlp;
/* max_batch bounds the sample count of sx_mlp_backward / sx_mlp_adam_step. */
int sx_mlp_create(sx_mlp** out, int64_t max_batch);
int sx_mlp_destroy(sx_mlp* mlp);
int sx_mlp_num_params(void);
int sx_mlp_set_params(sx_mlp* mlp, const float* h_params, void* stream);   /* resets Adam state */
int sx_mlp_get_params(sx_mlp* mlp, float* h_params, void* stream);
int sx_mlp_init(sx_mlp* mlp, uint64_t seed, void* stream);                 /* Glorot-uniform */
/* Features are normalised as (f - mu) / sigma; sx_mlp_predict_radius2 returns
 * exp(y + margin). */
int sx_mlp_set_normalizer(sx_mlp* mlp, const float* h_mu, const float* h_sigma, float margin, void* stream);
/* Raw prediction y (natural log of the squared k-th neighbour distance). */
int sx_mlp_forward(sx_mlp* mlp, const float* d_feats, int64_t n, float* d_out, void* stream);
int sx_mlp_predict_radius2(sx_mlp* mlp, const float* d_feats, int64_t n, float* d_radius2, void* stream);
/* Gradient of mean((y - target)^2) into d_grad (sx_mlp_num_params() floats,
 * device) and the summed squared error into *d_loss (device, nullable). */
int sx_mlp_backward(sx_mlp* mlp, const float* d_feats, const float* d_targets, int64_t n, float* d_grad,
                    float* d_loss, void* stream);
/* One Adam step (betas 0.9/0.999, eps 1e-8, decoupled weight decay wd). */
int sx_mlp_adam_step(sx_mlp* mlp, const float* d_feats, const float* d_targets, int64_t n, float lr, float wd,
                     float* d_loss, void* stream);
/* The 8 octree occupancy/depth features of index points [q_start, q_start+n_q)
 * into d_feats (n_q x 8 floats, 16-byte aligned). */
int sx_radius_features(sx_index* index, int64_t q_start, int64_t n_q, float* d_feats, void* stream);
/* Squared search radius exp(y + margin) of index points [q_start, q_start+n_q)
 * into d_radius2: the features and the MLP in one kernel (the feature rows
 * never reach HBM); equal to sx_radius_features + sx_mlp_predict_radius2. */
int sx_radius_predict(sx_index* index, sx_mlp* mlp, int64_t q_start, int64_t n_q, float* d_radius2, void* stream);

/* ---------------------------------------------------------------------------
 * Batched clouds (no counterpart in the reference, which builds one cloud at a
 * time: src/main.cu:80-107).  Cloud c is input points [offsets[c],
 * offsets[c+1]).  Every stage runs once over all clouds; per-cloud results
 * equal a standalone sx_index build of that cloud (textbook root).  Cloud c's
 * unique points occupy slots [cloud_base[c], cloud_base[c] + cloud_size[c])
 * of the per-slot arrays (cloud_base[c] a multiple of 64); neighbour rows and
 * normals are per slot, with indices relative to the cloud's first slot, and
 * rep[] is relative to the cloud's first input point.  With
 * sx_batch_set_octree, every build also makes each cloud's octree (the
 * reference's OT::Octree constructor, src/Octree.cu:213-319, once per cloud):
 * cloud c's nodes are oct_nodes[cloud_oct_base[c] ...] (cloud_oct_size[c]
 * nodes), byte-identical to a standalone index's octree of that cloud (child
 * links relative to the cloud), and the learned radius runs on the batch.
 * ------------------------------------------------------------------------- */
typedef struct sx_batch sx_batch;

typedef struct sx_batch_view {
    int32_t n_clouds;
    int64_t n_input;
    int64_t n_slots;            /* per-slot arrays hold n_slots entries       */
    int64_t n_unique;           /* unique points over all clouds              */
    const int64_t* cloud_offsets;  /* [n_clouds + 1] input offsets (device)   */
    const int32_t* cloud_base;     /* [n_clouds] first slot                   */
    const int32_t* cloud_size;     /* [n_clouds] unique points                */
    const float* cloud_bounds;     /* [n_clouds][4]: min, max, range, 0       */
    const uint64_t* codes;         /* [n_slots] unique Morton codes            */
    const uint32_t* rep;           /* [n_slots] first input index (cloud-local) */
    const float* points;           /* [n_slots][4] decoded points (x, y, z, 0) */
    const uint8_t* prefix_n;       /* [n_slots] radix tree, slot numbering     */
    const int32_t* left_child;
    const int32_t* parent;
    /* per-cloud octrees (NULL / 0 unless sx_batch_set_octree enabled them) */
    int64_t n_oct;                 /* octree nodes over all clouds            */
    const sx_otnode* oct_nodes;    /* [n_oct]                                 */
    const int32_t* cloud_oct_base; /* [n_clouds] first node of each cloud     */
    const int32_t* cloud_oct_size; /* [n_clouds] nodes of each cloud          */
    const uint8_t* has_leaf_left;  /* [n_slots] radix tree, slot numbering    */
    const uint8_t* has_leaf_right;
} sx_batch_view;

int sx_batch_create(sx_batch** out, int64_t max_points, int max_clouds, int code_bits);
int sx_batch_destroy(sx_batch* batch);
/* Host offsets [n_clouds + 1], offsets[0] = 0, non-decreasing; copied to the device (synchronous). */
int sx_batch_set_clouds(sx_batch* batch, const int64_t* h_offsets, int n_clouds, void* stream);
/* bucket: leaf-range size of the kNN; queue_blocks: deferred-query workgroups (0 = default). */
int sx_batch_set_tuning(sx_batch* batch, int bucket, int sort_algo, int queue_blocks);
/* oct_factor > 0: later builds (and graphs created later) also build every
 * cloud's octree, in storage for oct_factor x slot-capacity nodes allocated
 * here (3 is ample for scans; the view reports SX_ERANGE when a build needs
 * more); 0 frees it.  Synchronises the device; the batch must be rebuilt.
 * SX_ESTATE while graphs of the batch exist (they write into that storage). */
int sx_batch_set_octree(sx_batch* batch, float oct_factor);
/* Asynchronous, no host synchronisation, capturable. */
int sx_batch_build(sx_batch* batch, const float* d_x, const float* d_y, const float* d_z, void* stream);
/* Waits for the last sx_batch_build (whatever its stream); SX_EDATA when an
 * input coordinate is not finite. */
int sx_batch_get_view(sx_batch* batch, sx_batch_view* view, void* stream);
/* d_out: n_rows x k int32 (16-byte aligned); rows are written for slots
 * < min(n_rows, n_slots) (rows of padding slots are not written); n_rows is
 * the row capacity of the caller's buffers (n_slots from the view). */
int sx_batch_knn(sx_batch* batch, int k, int32_t* d_out, int64_t n_rows, void* stream);
/* kNN seeded by a per-slot squared search radius (the batch form of sx_knn in
 * SX_KNN_RADIUS mode): the same rows as sx_batch_knn, for any radii. */
int sx_batch_knn_radius(sx_batch* batch, int k, const float* d_radius2, int32_t* d_out, int64_t n_rows,
                        void* stream);
/* Learned-radius features (n_rows x 8 floats, 16-byte aligned) and fused
 * features + MLP radii of slots [0, n_rows); per cloud equal to
 * sx_radius_features / sx_radius_predict of a standalone index (rows of
 * padding slots are zero features).  Need the per-cloud octrees (SX_ESTATE). */
int sx_batch_radius_features(sx_batch* batch, float* d_feats, int64_t n_rows, void* stream);
int sx_batch_radius_predict(sx_batch* batch, sx_mlp* mlp, float* d_radius2, int64_t n_rows, void* stream);
int sx_batch_normals(sx_batch* batch, const int32_t* d_nn, int k, int precision, float vx, float vy, float vz,
                     float* d_normals, uint8_t* d_flags, int64_t n_rows, void* stream);
/* Queries the kNN tiles handed to the per-query search since the last reset. */
int sx_batch_deferred(sx_batch* batch, int64_t* deferred, int reset, void* stream);
/* One hipGraph of build + kNN (k) + FP32 normals (viewpoint at the origin) on
 * the given created stream; launching it reruns the whole batch. */
int sx_batch_graph_create(sx_batch* batch, const float* d_x, const float* d_y, const float* d_z, int k,
                          int32_t* d_nn, float* d_normals, int64_t n_rows, void** graph, void* stream);
/* SX_ESTATE when the graph's batch has been destroyed. */
int sx_batch_graph_launch(void* graph, void* stream);
int sx_batch_graph_destroy(void* graph);
/* Times the built batch's kNN for every (bucket, queue_blocks) candidate and
 * keeps the fastest.  h_ms (nullable): n_buckets x n_queue times in ms. */
int sx_batch_autotune(sx_batch* batch, int k, int32_t* d_scratch_nn, int64_t n_rows, const int* h_buckets, int n_buckets,
                      const int* h_queue_blocks, int n_queue, int reps, float* h_ms, int* best_bucket,
                      int* best_queue_blocks, void* stream);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* SHAPES_C_H_ */
