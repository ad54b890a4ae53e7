// dem_verlet.inl -- candidate lists with a skin for pure-DEM steps.
// Included by dem_hip.hip after the contact-scan kernels.
//
// The reference re-sorts the grid and scans 27 cells for every component on
// every step (contact:1121-1143).  Here the grid is rebuilt only when some
// component centre has moved more than skin/2 since the last rebuild.  At a
// rebuild every component lists the partners within r_i + r_j + skin
// (scanning +-2 cells, which covers that radius for cells of at least
// 2 r_max); in between, every listed partner is re-tested with the exact
// overlap test of the full scan, restricted to the 27 cells around the
// component's current cell as the reference is, and the contacts are put in
// the reference's candidate order: cell offset (dx, dy, dz) lexicographic,
// then input row within a cell.  Any pair that overlaps now was within
// r_i + r_j + skin at the rebuild (each centre moved at most skin/2), so the
// contact set, its order and hence every slot and force are the full scan's.

constexpr int CAND_MAX = 96;  // listed partners per component

// Rebuild: partners within r_i + r_j + skin (rows; the grid's cell ranges
// are sorted positions, mapped to rows by `sorted`), and the centres the
// displacement test compares against.  Cells whose box lies farther from the
// centre than r_i + the scene's largest radius + skin are skipped (whole
// columns, and the ends of a column's z run, so each column stays one sorted
// range in the same order): with cells of 2 r_max the +-2 ring is out of
// reach unless the centre is within a skin of its own cell's face.
__global__ void __launch_bounds__(BLOCK) k_vbuild(Soa S, const double4* posr, const unsigned long long* key,
                                                  const int* sorted, const unsigned* start, const SceneGrid* grid,
                                                  const SceneParams* sp, Scalars* sc, double skin, int* cand,
                                                  int* ncand, double* refC)
{
    const long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const long n = S.n;
    if (t >= n) return;
    const double4 me = posr[t];
    refC[t] = me.x;
    refC[n + t] = me.y;
    refC[2 * n + t] = me.z;
    int count = 0;
    if (S.k(I_VALID, t)) {
        const int scene = S.k(I_SCENE, t);
        const SceneGrid& g = grid[scene];
        const double h = 1.0 / sp[scene].inv;
        const double far = me.w + sp[scene].maxr + skin + cell_margin(me, h);
        const double far2 = far * far;
        const long X = (long)g.dims[0], Y = (long)g.dims[1], Z = (long)g.dims[2];
        const long ax = S.k(I_CX, t), ay = S.k(I_CY, t), az = S.k(I_CZ, t);
        const long cx = ax - g.lo[0], cy = ay - g.lo[1], cz = az - g.lo[2];
        const bool dense = sc->dense;
        for (long dx = -2; dx <= 2; ++dx) {
            const long xx = cx + dx;
            if (xx < 0 || xx >= X) continue;
            const double gx = slab_gap(me.x, ax + dx, h);
            if (gx >= far) continue;
            for (long dy = -2; dy <= 2; ++dy) {
                const long yy = cy + dy;
                if (yy < 0 || yy >= Y) continue;
                const double gy = slab_gap(me.y, ay + dy, h);
                const double gxy = gx * gx + gy * gy;
                if (gxy >= far2) continue;
                long zlo = cz > 1 ? cz - 2 : 0, zhi = cz + 2 < Z ? cz + 2 : Z - 1;
                while (zlo < cz) {
                    const double gz = slab_gap(me.z, zlo + g.lo[2], h);
                    if (gxy + gz * gz < far2) break;
                    ++zlo;
                }
                while (zhi > cz) {
                    const double gz = slab_gap(me.z, zhi + g.lo[2], h);
                    if (gxy + gz * gz < far2) break;
                    --zhi;
                }
                const unsigned long long col = g.base + ((unsigned long long)xx * Y + yy) * Z;
                long b, e;
                if (dense) {
                    b = start[col + zlo];
                    e = start[col + zhi + 1];
                } else {
                    b = lower_bound_u64(key, n, col + zlo);
                    e = lower_bound_u64(key, n, col + zhi + 1);
                }
                for (long k = b; k < e; ++k) {
                    const long q = sorted[k];
                    if (q == t) continue;
                    const double4 c = posr[q];
                    const double ex = c.x - me.x, ey = c.y - me.y, ez = c.z - me.z;
                    const double reach = me.w + c.w + skin;
                    if (ex * ex + ey * ey + ez * ez >= reach * reach) continue;
                    if (count < CAND_MAX) cand[(long)count * n + t] = (int)q;
                    ++count;
                }
            }
        }
    }
    if (count > CAND_MAX) atomicAdd(&sc->cand_overflow, 1);
    ncand[t] = min(count, CAND_MAX);
}

// The reference's candidate order of a contact of a component in packed
// cell `mine` (pack_meta) with the component whose record is `other`: cell
// offset (dx, dy, dz) lexicographic, then input row; ~0u when `other` lies
// outside the 27 cells around `mine` (no contact).
__device__ __forceinline__ uint32_t order_key(unsigned long long mine, int4 other)
{
    constexpr unsigned long long M = (1ull << 21) - 1;
    const unsigned long long c = meta_cell(other);
    const int ddx = (int)((c >> 42) & M) - (int)((mine >> 42) & M);
    const int ddy = (int)((c >> 21) & M) - (int)((mine >> 21) & M);
    const int ddz = (int)(c & M) - (int)(mine & M);
    if ((c >> 63) || ddx < -1 || ddx > 1 || ddy < -1 || ddy > 1 || ddz < -1 || ddz > 1) return ~0u;
    return ((uint32_t)(((ddx + 1) * 3 + (ddy + 1)) * 3 + (ddz + 1)) << 25) | (uint32_t)other.z;
}

// List entries of a row during the list scan: (partner row << SLOT_BITS |
// slot) for a contact that holds a history slot already, ~(partner row) for a
// new contact, which takes its slot after the scan.
__device__ __forceinline__ long entry_row(int e) { return e >= 0 ? (long)(e >> SLOT_BITS) : (long)~e; }

// The new contacts of a row's list take their slots in list order
// (acquireHistorySlot, contact:190-215: lookups of existing contacts do not
// change the occupancy, so taking the new ones after the others is the same
// sequence of allocations); a contact left without one (all 64 taken) is
// dropped, as the reference drops it.  Returns the entries kept.
__device__ __forceinline__ int assign_new(long t, long n, int found, const int4* meta, unsigned long long& occ,
                                          unsigned long long& cur, uint32_t (&k16)[16], uint32_t* keys,
                                          double* vals, long vstride, int* nbr, int& high, int& over)
{
    int kept = 0;
    for (int c = 0; c < found; ++c) {
        int e = nbr[(long)c * n + t];
        if (e < 0) {
            const long k = ~e;
            const int slot = new_pp(occ, cur, k16, keys, vals, vstride, (uint32_t)meta[k].w + 1u, high, over);
            if (slot < 0) continue;  // slot overflow: counted, no force
            e = ((int)k << SLOT_BITS) | slot;
        }
        nbr[(long)kept * n + t] = e;
        ++kept;
    }
    return kept;
}

// Step between rebuilds: exact contact test over the listed partners, the
// reference's candidate order, then history slots.  Rows are their own
// sorted positions on these steps.  Per batch of VSCAN_BATCH candidates the
// indices, then the centres, then (for the overlapping ones) the cell
// records with the component ids are fetched together; the row's history
// state is fetched before the scan.  A contact that holds a slot finds it as
// the scan meets it (the lookup does not depend on the order); new contacts
// take theirs after the scan, in list order (assign_new).
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5))) k_vscan(Soa S, const double4* posr, const int4* meta,
                                                 const uint32_t* cidS, const int* cand, const int* ncand,
                                                 Scalars* sc, History H, int step, int* nbr,
                                                 int* ncount, unsigned long long* occTape, bool counters,
                                                 unsigned* bsum, const int* skip, int* mirror, int seq)
{
    if (mirror && blockIdx.x == 0 && threadIdx.x == 0) {
        // the rebuild flag and range count of k_cells, for the host (which
        // waits for `seq`)
        mirror[1] = sc->need_rebuild;
        mirror[2] = sc->grid_overflow;
        __hip_atomic_store(mirror, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    if (skip && *skip) return;  // launched ahead of a decision that came out as a rebuild
    const long b = xcd_block(blockIdx.x, gridDim.x);
    const long t = b * BLOCK + threadIdx.x;
    const long n = S.n;
    int found = 0, stored = 0, high = 0, over = 0, fresh = 0;
    uint32_t last = 0;  // order key of the last contact in the list
    if (t < n) {
        // history state of the row (expiry: slots not seen last step drop);
        // on list steps every row's history is its own row (rebuilds move
        // the histories with the rows; differentiable runs move no row)
        const long o = t;
        unsigned long long occ = H.ppOcc[o] & H.ppSeen[(long)((step + 1) & 1) * n + o];
        unsigned long long cur = 0;  // slots seen this step
        uint32_t* keys = H.ppKey + o * PP_SLOTS;
        uint32_t k16[16];
        {
            const uint4* kv = reinterpret_cast<const uint4*>(keys);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint4 v = kv[i];
                k16[4 * i] = v.x; k16[4 * i + 1] = v.y; k16[4 * i + 2] = v.z; k16[4 * i + 3] = v.w;
            }
        }
        const unsigned long long mine = meta_cell(meta[t]);
        if (!(mine >> 63)) {  // inside the hash range
            const double4 me = posr[t];
            const int m = ncand[t];
            constexpr int VSCAN_BATCH = 4;
            for (int c0 = 0; c0 < m; c0 += VSCAN_BATCH) {
                long kq[VSCAN_BATCH];
                double4 oq[VSCAN_BATCH];
                bool hit[VSCAN_BATCH];
#pragma unroll
                for (int j = 0; j < VSCAN_BATCH; ++j) kq[j] = c0 + j < m ? cand[(long)(c0 + j) * n + t] : t;
#pragma unroll
                for (int j = 0; j < VSCAN_BATCH; ++j) oq[j] = posr[kq[j]];
#pragma unroll
                for (int j = 0; j < VSCAN_BATCH; ++j) {
                    const double4 c = oq[j];
                    const double dx_ = c.x - me.x, dy_ = c.y - me.y, dz_ = c.z - me.z;
                    const double d2 = dx_ * dx_ + dy_ * dy_ + dz_ * dz_;
                    const double rs = me.w + c.w;
                    hit[j] = c0 + j < m && !(d2 >= rs * rs || d2 <= 1.0e-28);
                }
                // cell records (with the component ids) of the overlapping
                // candidates (the others read the row's own, which it read
                // already)
                int4 mq[VSCAN_BATCH];
#pragma unroll
                for (int j = 0; j < VSCAN_BATCH; ++j) mq[j] = meta[hit[j] ? kq[j] : t];
#pragma unroll
                for (int j = 0; j < VSCAN_BATCH; ++j) {
                    if (!hit[j]) continue;
                    const long k = kq[j];
                    const uint32_t kk = order_key(mine, mq[j]);
                    if (kk == ~0u) continue;  // beyond the 27 cells
                    if (found >= MAX_LIST) {
                        ++found;
                        continue;
                    }
                    const int slot = find_pp(occ, k16, keys, (uint32_t)mq[j].w + 1u);
                    int e = ~(int)k;  // a new contact
                    if (slot >= 0) {
                        e = ((int)k << SLOT_BITS) | slot;
                        cur |= 1ull << slot;
                        high = max(high, slot + 1);
                    } else {
                        ++fresh;
                    }
                    // the contacts in order: candidate lists keep the order of
                    // the rebuild, so a contact almost always comes after the
                    // last one (kept in a register); otherwise it is inserted,
                    // the keys of the entries it passes recomputed from their
                    // cell records
                    if (found == 0 || kk > last) {
                        nbr[(long)found * n + t] = e;
                        last = kk;
                    } else {
                        int pos = found;
                        while (pos > 0) {
                            const int ep = nbr[(long)(pos - 1) * n + t];
                            if (order_key(mine, meta[entry_row(ep)]) <= kk) break;
                            nbr[(long)pos * n + t] = ep;
                            --pos;
                        }
                        nbr[(long)pos * n + t] = e;
                    }
                    ++found;
                }
            }
        }
        if (found > MAX_LIST) {
            // more contacts than a list holds: visit them in the reference's
            // order by selection (the next larger order key each time) while
            // the slots are acquired
            const double4 me = posr[t];
            const int m = ncand[t];
            const int total = found;
            stored = assign_inline(S, t, t, ident_order(), cidS, H, step, nbr, occTape, high, over,
                                   [&](auto&& acq) {
                uint32_t last = 0;
                for (int r = 0; r < total; ++r) {
                    uint32_t best = 0xFFFFFFFFu;
                    long bk = -1;
                    for (int c = 0; c < m; ++c) {
                        const long k = cand[(long)c * n + t];
                        const double4 o = posr[k];
                        const double dx_ = o.x - me.x, dy_ = o.y - me.y, dz_ = o.z - me.z;
                        const double d2 = dx_ * dx_ + dy_ * dy_ + dz_ * dz_;
                        const double rs = me.w + o.w;
                        if (d2 >= rs * rs || d2 <= 1.0e-28) continue;
                        const uint32_t kk = order_key(mine, meta[k]);
                        if (kk == ~0u) continue;  // beyond the 27 cells
                        if ((r == 0 || kk > last) && kk < best) {
                            best = kk;
                            bk = k;
                        }
                    }
                    if (bk < 0) break;
                    acq(bk);
                    last = best;
                }
            });
        } else {
            if (occTape) occTape[o] = occ;
            stored = fresh ? assign_new(t, n, found, meta, occ, cur, k16, keys, pp_slot(H, o, 0), H.rows * 9, nbr,
                                        high, over)
                           : found;
            H.ppOcc[o] = occ;
            H.ppSeen[(long)(step & 1) * n + o] = cur;
        }
        ncount[t] = stored;
    }
    block_count(stored, high, bsum, b);
    atomic_add_wave(&sc->pp_overflow, over);
}

// History rows follow the particles: after a rebuild reordered the particle
// arrays, row t's history moves from its old row (I_HIST) to row t, so the
// per-contact history traffic of the following steps is as local as the
// particle data.  Only the slots below the block's highest occupied one are
// copied (a free slot is never read before acquisition resets it, so the
// free slots of a row below that bound may be copied too).  The pair slots
// move as whole tiles: for each slot, the block's rows are one contiguous
// run of 9-f64 values in the destination, copied one f64 per lane, and the
// rows' keys move in 16-byte quads.
__global__ void __launch_bounds__(BLOCK) k_hist_move(Soa S, History src, History dst)
{
    __shared__ long hrow[BLOCK];
    __shared__ int topw[BLOCK / 64];
    const long t0 = (long)blockIdx.x * BLOCK;
    const long t = t0 + threadIdx.x;
    const long n = S.n;
    const int rowsHere = (int)min((long)BLOCK, n - t0);
    int top = 0;
    if (t < n) {
        const long h = S.k(I_HIST, t);
        hrow[threadIdx.x] = h;
        const unsigned long long occ = src.ppOcc[h];
        dst.ppOcc[t] = occ;
        dst.ppSeen[t] = src.ppSeen[h];
        dst.ppSeen[n + t] = src.ppSeen[n + h];
        top = occ ? 64 - __clzll((long long)occ) : 0;
    }
    top = wave_max(top);
    if ((threadIdx.x & 63) == 0) topw[threadIdx.x >> 6] = top;
    __syncthreads();
    int bt = topw[0];
    for (int w = 1; w < BLOCK / 64; ++w) bt = max(bt, topw[w]);
    const int q4 = (bt + 3) >> 2;  // key quads per row
    for (int e = threadIdx.x; e < rowsHere * q4; e += BLOCK) {
        const int r = e / q4, q = e - r * q4;
        reinterpret_cast<uint4*>(dst.ppKey + (t0 + r) * PP_SLOTS)[q] =
            reinterpret_cast<const uint4*>(src.ppKey + hrow[r] * PP_SLOTS)[q];
    }
    for (int b = 0; b < bt; ++b) {
        const double* sv = pp_slot(src, 0, b);
        double* dv = pp_slot(dst, t0, b);
        for (int e = threadIdx.x; e < rowsHere * 9; e += BLOCK) {
            const int r = e / 9, k = e - r * 9;
            dv[e] = sv[hrow[r] * 9 + k];
        }
    }
    if (t >= n) return;
    const long h = hrow[threadIdx.x];
    const unsigned wocc = src.pwOcc[h];
    dst.pwOcc[t] = wocc;
    dst.pwSeen[t] = src.pwSeen[h];
    dst.pwSeen[n + t] = src.pwSeen[n + h];
    const int wtop = wocc ? 32 - __clz((int)wocc) : 0;
    for (int b = 0; b < wtop; ++b) dst.pwKey[t * PW_SLOTS + b] = src.pwKey[h * PW_SLOTS + b];
    const double* sw = src.pwVal + h * PW_SLOTS * 9;
    double* dw = dst.pwVal + t * PW_SLOTS * 9;
    for (int k = 0; k < wtop * 9; ++k) dw[k] = sw[k];
    if (src.ptOcc) {  // triangle slots (runs with bodies)
        const unsigned long long bocc = src.ptOcc[h];
        dst.ptOcc[t] = bocc;
        dst.ptSeen[t] = src.ptSeen[h];
        dst.ptSeen[n + t] = src.ptSeen[n + h];
        const int btop = bocc ? 64 - __clzll((long long)bocc) : 0;
        for (int b = 0; b < btop; ++b) dst.ptKey[t * PT_SLOTS + b] = src.ptKey[h * PT_SLOTS + b];
        const double* sb = src.ptVal + h * PT_SLOTS * 9;
        double* db = dst.ptVal + t * PT_SLOTS * 9;
        for (int k = 0; k < btop * 9; ++k) db[k] = sb[k];
    }
    S.k(I_HIST, t) = (int)t;
}
