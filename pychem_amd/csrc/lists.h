// lists.h -- screened quartet lists built on the device, and the ket-side J
// reduction of the tiled J/K digestion.
//
// Pairs of one pair class are sorted by (molecule, Schwarz bound descending).
// For a bra pair i the kept kets are then a contiguous run of the sorted ket
// list: all kets j of the molecule (j >= i when bra and ket class agree) with
// bound_i * bound_j > thr, the diagonal (i|i) always included.  Quartets with
// bound product below thr32 form a second (FP32) list; since the product falls
// monotonically along the ket list, each bra's FP64 kets are a prefix of its
// run and its FP32 kets the rest.  The same monotonicity in the bra index
// makes the bras that reach a given ket j a contiguous range, which is what
// the ket J reduction walks.
#pragma once
#include <vector>

#include "jk_common.h"

struct ListArgs {
    const int* bra;              // bra pair ids sorted by (molecule, bound desc)
    const int* ket;              // ket pair ids, same order (the same list when same != 0)
    const int* bra_seg;          // [nmol + 1] molecule segment starts in bra
    const int* ket_seg;          // [nmol + 1] molecule segment starts in ket
    const long long* ket_npcum;  // [nket + 1] cumulative primitive-pair counts along ket
    const double* bound;         // Schwarz bound per pair id
    const double* bra_bound;     // bound of bra[i], in bra order
    const PairHdr* pairs;
    const int* cart_off;
    double thr, thr32;
    int nb, nk, same;
};

// Per bra of one list (FP64: w = 0, FP32: w = 1).
struct ListBras {
    int* j0;          // first ket index in the sorted ket list
    long long* cnt;   // kets kept
    long long* np;    // primitive quartets
    long long* qs;    // first quartet (exclusive scan of cnt)
    long long* nit;   // J/K work items
    long long* its;   // first work item (exclusive scan of nit)
};

// Host launchers (stream-ordered; no synchronisation).
void list_count(const ListArgs& a, const ListBras& b64, const ListBras& b32, hipStream_t st);
// Exclusive scan of n values (in may equal out); *total = sum.
void list_scan(const long long* in, long long* out, int n, long long* total, hipStream_t st);
// J/K work items of a class (nit / its of b64): per bra, its FP64 kets then
// its FP32 kets (b32, may be NULL) in chunks of `run` quartets; items32 gets
// each item's FP32 range.
void list_items_count(const ListBras& b64, const ListBras* b32, int nb, long long run, hipStream_t st);
void list_fill(const ListArgs& a, const ListBras& b, long long nq, int2* quartets, unsigned* kets, hipStream_t st);
void list_items_fill(const ListArgs& a, const ListBras& b64, const ListBras* b32, long long nitems, long long run,
                     JkItem* items, int2* items32, hipStream_t st);
// Batched construction: one launch per phase over every class of a plan.
// A job is one ERI class: its bra/ket lists and both (FP64, FP32) quartet
// lists.  Work is laid out job after job in whole blocks of 256 threads;
// block0[j] is job j's first block (block0[n] = all blocks), so the job of a
// block is found with scalar loads from the kernel arguments.
#define LIST_MAXJOB 48
struct ListJob {
    ListArgs a;            // a.nb = 0: nothing to count for this class
    ListBras b64, b32;
    int2* q[2];            // quartet lists (FP64, FP32)
    unsigned* k[2];        // ket records
    long long nq[2];
    JkItem* items;
    int2* items32;         // NULL without an FP32 list
    long long nitems;
    long long run;         // work-item length
    int nb_items;          // bras whose items are counted (0: class without quartets)
    int has32;             // the class has an FP32 list
    int* ord[2];           // quartet order of the ERI kernels (list_order_batch; NULL: none)
};
struct ListBlocks {
    int block0[LIST_MAXJOB + 1];
    int n;
};
struct ScanJob {
    const long long* in;
    long long* out;
    int n;
    long long* total;
};
// per bra of every job: list_count_kernel's runs (block0 over jobs' a.nb)
void list_count_batch(const ListJob* jobs, const ListBlocks& blk, hipStream_t st);
// one single-workgroup exclusive scan per ScanJob (jobs in device memory)
void list_scan_batch(const ScanJob* jobs, int n, hipStream_t st);
// every quartet of every list: jobs' list w = 0, 1 as blocks 2j + w
void list_fill_batch(const ListJob* jobs, const ListBlocks& blk, hipStream_t st);
// per bra of every job (block0 over nb_items): work items of its kets
void list_items_count_batch(const ListJob* jobs, const ListBlocks& blk, hipStream_t st);
// Order in which the ERI kernels visit each list (jobs' q[w], ord[w]; the
// blocks of list_fill_batch): quartets by descending primitive work
// (primitive pairs of bra x ket, in half-octave buckets), so that the lanes
// of a wave loop over about as many primitive quartets as one another.  A
// counting sort: per-bucket counts, their scan, then each quartet's slot.
// cnt and cur: ORD_NB ints per list (2 per job); cnt zeroed here.  blk:
// ceil(nq / (256 ORD_PT)) blocks per list, none for lists without an order.
#define ORD_NB 32
#define ORD_PT 16
void list_order_batch(const ListJob* jobs, const ListBlocks& blk, int* cnt, int* cur, hipStream_t st);
// every work item of every job (block0 over nitems)
void list_items_fill_batch(const ListJob* jobs, const ListBlocks& blk, hipStream_t st);

// J'_Q += sum over the bras that reach ket Q of the per-quartet ket J values
// jkbuf[(tile * ncd + cd) * nq + q], over tiles of up to 64 kets x 256 bras of
// one molecule (int4: first ket, kets, first bra, bras); one list per launch,
// or every list of a table in one launch.
struct JredList {
    ListArgs a;
    ListBras b;
    const double* jkbuf;
    long long nq;
    int w, ntiles, ncd, n4;  // ntiles: bra component tiles of the digestion kernel
    const int4* tiles;
    int ntile;
    long long block0;        // first block of this list in the fused launch (ntile * ncd blocks)
};
std::vector<int4> jred_tiles(const std::vector<int>& bra_seg, const std::vector<int>& ket_seg,
                             const std::vector<double>& bra_bound, const std::vector<double>& ket_bound, int same,
                             double thr);
JredList jred_list(const ListArgs& a, const ListBras& b, int w, const double* jkbuf, long long nq, int ntiles, int ncd,
                   int n4, const int4* tiles, int ntile);
void jket_reduce_all(const JredList* lists_dev, int nlist, long long nblocks, const MolInfo* mols, double* J,
                     hipStream_t st);
void jket_reduce(const JredList& L, const MolInfo* mols, double* J, hipStream_t st);
