// Job tables and work lists of batched transforms.
//
// Each image contributes a row job and a column job, cut into tiles of whole
// lines.  A small image whose rows and columns each fit one tile becomes a
// single "whole" item done by one workgroup.  Without dependencies, items
// are sorted by phase (whole images and row tiles first, column tiles
// second) and then by kernel variant (padded length, Bluestein or not, axis)
// so neighbouring workgroups run the same code.  A line length gets the same
// radix schedule in every batch, so a result is bitwise the same alone or in
// any mix.
#include "batch_builder.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <tuple>

#include "../../../include/mdft.h"

namespace mdft {

void arrange_for_xcds(const std::vector<uint2>& items, const std::vector<double>& cost, std::vector<uint2>& out);

int batch_threads_for(const uint32_t* widths, const uint32_t* heights, size_t count)
{
    const int limit = batch_max_logm(1024);
    int tpl = 1;
    for (size_t i = 0; i < count; ++i) {
        compute_m(widths[i]);  // validates the sides
        compute_m(heights[i]);
        // the FFT lengths the tiles run (the padded length, or 3 * 2^k)
        const int sr = line_shape(widths[i], false), sc = line_shape(heights[i], true);
        if (m_of(sr) > (1 << limit) || m_of(sc) > (1 << limit))
            throw Error(MDFT_ERR_SIZE, "batch transform " + std::to_string(i) + " (" + std::to_string(widths[i]) + "x" +
                                           std::to_string(heights[i]) + ") exceeds the batch padded-length limit " +
                                           std::to_string(1 << limit));
        tpl = std::max({tpl, m_of(sr) / ept_of(sr), m_of(sc) / ept_of(sc)});
    }
    // the smallest tile that holds a whole line of every length
    return tpl <= 256 ? 256 : tpl <= 512 ? 512 : 1024;
}

void build_batch(Context& ctx, const BatchEntry* entries, size_t count, int T, int TC, bool dependent,
                 BatchBuild& out, std::vector<uint32_t>* job_image, std::vector<uint8_t>* job_is_col)
{
    out = BatchBuild{};
    out.threads = T;
    out.threads_cols = TC;
    if (job_image) job_image->clear();
    if (job_is_col) job_is_col->clear();

    std::vector<uint32_t> order(count);
    std::iota(order.begin(), order.end(), 0u);
    std::vector<double> work(count);
    for (size_t i = 0; i < count; ++i)
        work[i] = (double)compute_m(entries[i].width) * entries[i].height +
                  (double)compute_m(entries[i].height) * entries[i].width;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return work[a] > work[c]; });

    struct Item {
        int phase;
        int m, shape, blue, kind;  // kind: 0 whole, 1 row tile, 2 column tile
        uint32_t seq;
        uint2 v;
    };
    std::vector<Item> items;
    std::vector<std::pair<std::vector<uint2>, std::vector<uint2>>> per_image;  // dependency ordering
    uint32_t counter = 0, seq = 0;
    for (uint32_t idx : order) {
        const BatchEntry& e = entries[idx];
        const int w = (int)e.width, h = (int)e.height;
        auto rp = ctx.line_plan(w, false);
        auto cp = ctx.line_plan(h, true);
        out.keep.push_back(rp);
        out.keep.push_back(cp);
        const int lpw_r = w > 1 ? batch_lines_per_tile(T, rp->shape) : 0;
        // a whole image runs both passes in one row-phase workgroup
        const int lpw_cw = h > 1 ? batch_lines_per_tile(T, cp->shape) : 0;
        const bool whole = (w > 1 && h > 1 && h <= lpw_r && w <= lpw_cw);
        // column tiles of images with a row pass run in the column phase (TC
        // threads); a single-column-pass image (w == 1) runs with the rows
        const int lpw_c = h > 1 ? batch_lines_per_tile((w > 1 && !whole) ? TC : T, cp->shape) : 0;
        const int ntr = w > 1 ? (h + lpw_r - 1) / lpw_r : 0;
        const int ntc = h > 1 ? (w + lpw_c - 1) / lpw_c : 0;
        const int sig = (dependent && w > 1 && h > 1 && !whole) ? (int)counter++ : -1;
        int rj = -1, cj = -1;
        if (w > 1) {
            BatchJob j{};
            j.job = ctx.make_job(*rp, e.src, e.dst, 1, w, h, e.inverse);
            j.shape = rp->shape;
            j.blue = rp->blue;
            j.cols = 0;
            j.ntiles = ntr;
            j.dep = -1;
            j.sig = sig;
            rj = (int)out.jobs.size();
            out.jobs.push_back(j);
            if (job_image) job_image->push_back(idx);
            if (job_is_col) job_is_col->push_back(0);
        }
        if (h > 1) {
            BatchJob j{};
            // the column pass reads the row pass's output when there is one
            j.job = ctx.make_job(*cp, w > 1 ? e.dst : e.src, e.dst, w, 1, w, e.inverse);
            j.shape = cp->shape;
            j.blue = cp->blue;
            j.cols = 1;
            j.ntiles = ntc;
            j.dep = sig;
            j.dep_need = ntr;
            j.sig = -1;
            cj = (int)out.jobs.size();
            out.jobs.push_back(j);
            if (job_image) job_image->push_back(idx);
            if (job_is_col) job_is_col->push_back(1);
        }
        std::vector<uint2> ri, ci;
        if (whole) {
            items.push_back({1, rp->m, rp->shape, rp->blue, 0, seq++, make_uint2(rj, BATCH_WHOLE | (uint32_t)cj)});
            ri.push_back(items.back().v);
        } else {
            for (int t = 0; t < ntr; ++t) {
                items.push_back({1, rp->m, rp->shape, rp->blue, 1, seq++, make_uint2(rj, t)});
                ri.push_back(items.back().v);
            }
            for (int t = 0; t < ntc; ++t) {
                items.push_back({w > 1 ? 2 : 1, cp->m, cp->shape, cp->blue, 2, seq++, make_uint2(cj, t)});
                ci.push_back(items.back().v);
            }
        }
        per_image.push_back({ri, ci});
    }

    if (dependent) {
        // rows of image k+1 before columns of image k: a column tile only
        // waits on row tiles with smaller tickets
        for (size_t k = 0; k < per_image.size(); ++k) {
            out.items.insert(out.items.end(), per_image[k].first.begin(), per_image[k].first.end());
            if (k > 0)
                out.items.insert(out.items.end(), per_image[k - 1].second.begin(), per_image[k - 1].second.end());
        }
        if (!per_image.empty())
            out.items.insert(out.items.end(), per_image.back().second.begin(), per_image.back().second.end());
        out.nphase1 = (uint32_t)out.items.size();
    } else {
        std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& c) {
            return std::make_tuple(a.phase, -a.m, a.shape, a.blue, a.kind, a.seq) <
                   std::make_tuple(c.phase, -c.m, c.shape, c.blue, c.kind, c.seq);
        });
        std::vector<uint2> p1, p2;
        std::vector<double> c1, c2;
        for (const Item& it : items) {
            // cost of a tile ~ points x passes (x2 for Bluestein's two FFTs)
            const int lpw = batch_lines_per_tile(it.kind == 2 && it.phase == 2 ? TC : T, it.shape);
            const double cost = double(lpw) * double(it.m) * (std::log2(double(it.m)) + 2) * (it.blue ? 2.0 : 1.0) *
                                (it.kind == 0 ? 2.0 : 1.0);
            (it.phase == 1 ? p1 : p2).push_back(it.v);
            (it.phase == 1 ? c1 : c2).push_back(cost);
        }
        out.nphase1 = (uint32_t)p1.size();
        arrange_for_xcds(p1, c1, out.items);
        arrange_for_xcds(p2, c2, out.items);
    }
    out.ncounters = counter;
}

// Workgroups of a launch are dealt to the XCDs round-robin: block b runs on
// XCD b % 8, and each XCD has its own L2.  Give every job (its tiles, and
// with them its chirp, diagonal and twiddle tables and, for column tiles,
// the cache lines neighbouring tiles share) to one XCD, balancing the XCDs'
// total cost (largest job first, jobs larger than an eighth of the launch
// split), then lay the eight queues out so that queue x fills the blocks
// b % 8 == x.  Within a queue the variant order of `items` is kept.
void arrange_for_xcds(const std::vector<uint2>& items, const std::vector<double>& cost, std::vector<uint2>& out)
{
    constexpr int NXCD = 8;
    const size_t n = items.size();
    if (n == 0) return;
    double total = 0;
    for (double c : cost) total += c;
    // runs of consecutive items of one job (items are grouped by variant, then job)
    struct Run {
        size_t begin, end;
        double cost;
    };
    std::vector<Run> runs;
    const double cap = total / (2 * NXCD);
    for (size_t i = 0; i < n;) {
        size_t j = i;
        double c = 0;
        while (j < n && items[j].x == items[i].x && (j == i || c + cost[j] <= cap)) c += cost[j++];
        runs.push_back({i, j, c});
        i = j;
    }
    std::vector<size_t> by_cost(runs.size());
    std::iota(by_cost.begin(), by_cost.end(), size_t(0));
    std::stable_sort(by_cost.begin(), by_cost.end(), [&](size_t a, size_t b) { return runs[a].cost > runs[b].cost; });
    std::vector<int> owner(runs.size());
    double load[NXCD] = {};
    for (size_t r : by_cost) {
        int best = 0;
        for (int x = 1; x < NXCD; ++x)
            if (load[x] < load[best]) best = x;
        owner[r] = best;
        load[best] += runs[r].cost;
    }
    std::vector<uint2> queue[NXCD];
    for (size_t r = 0; r < runs.size(); ++r)
        for (size_t i = runs[r].begin; i < runs[r].end; ++i) queue[owner[r]].push_back(items[i]);
    // interleave; when a queue runs dry its slots take the longest remaining queue's items
    size_t pos[NXCD] = {};
    for (size_t b = 0; b < n; ++b) {
        int x = int(b % NXCD);  // block index within this launch
        if (pos[x] == queue[x].size()) {
            int best = -1;
            for (int y = 0; y < NXCD; ++y)
                if (pos[y] < queue[y].size() && (best < 0 || queue[y].size() - pos[y] > queue[best].size() - pos[best]))
                    best = y;
            x = best;
        }
        out.push_back(queue[x][pos[x]++]);
    }
}

}  // namespace mdft
