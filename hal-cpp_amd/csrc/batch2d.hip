// batch2d.hip -- many independent ADI wells of uneven size advanced together.
//
// Reference: the host-side runner MultiWellExperiment::Step / StepMultiThread
// (include/Hal/Tools/MultiWellExperiment.h:82-103), which steps one model per
// well, serially or on a thread pool.  Here every well is a HAL grid ticked as
// DiffusionADI(rate_w) + MulAll(-decay_w) + Add(sources_w) + Update
// (include/Hal/Grids/PDEGrid2D.h:103-218, 328-347), and a tick of the whole
// batch is two launches over a work list covering every well:
//   BatchScan   one block per (well, 16-column group): the cross-strip scan
//   BatchStrip  one block per (well, strip): sweep A finish, sweep B, decay,
//               point sources and the next tick's strip reduction.  Wells are
//               grouped by width (<= 256, <= 512, <= 1024 columns) and each
//               width class has a launch of its own with its own workgroup
//               shape (threads x columns per thread), so a class's register
//               budget and occupancy are its own; vector or scalar rows are a
//               block-uniform switch.
// All are captured in one hipGraph per chunk of ticks.  The shape of every
// width class, the strip height and the carry method are picked by an
// autotuner (timed candidates, cached per batch signature).  The per-instance
// launch loop (halhip_batch_run_ticks_per_instance) runs the same kernels one
// well at a time and is the baseline the batch is measured against.
#include "adi_device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <string>

using namespace halhip;

namespace
{
    constexpr int kBatchMaxSide = 1024;
    constexpr int kClasses = 3;       // well widths <= 256, <= 512, <= 1024
    constexpr int kShapesPerClass = 3;

    // Workgroup shapes (threads, columns per thread) of each width class;
    // threads * columns covers the class's widest well.  Option 0 is the
    // default.
    struct BlockShape
    {
        int32_t threads;
        int32_t cpt;
    };
    constexpr BlockShape kShapes[kClasses][kShapesPerClass] = {
        {{256, 1}, {128, 2}, {64, 4}},
        {{256, 2}, {512, 1}, {128, 4}},
        {{256, 4}, {512, 2}, {1024, 1}},
    };

    int32_t WidthClass(int32_t Y)
    {
        return (Y <= 256) ? 0 : ((Y <= 512) ? 1 : 2);
    }

    struct BatchWell
    {
        AdiArgs args;
        int64_t fieldOffset;
        int32_t vec; // rows move as vectors (yDim a multiple of the columns per thread)
        int32_t pad;
        float *redHalf[2]; // windowed carries: the two halves of this well's reductions
    };

    struct StripItem
    {
        int32_t well;
        int32_t strip;
        int32_t srcBegin;
        int32_t srcEnd;
    };

    struct ScanItem
    {
        int32_t well;
        int32_t yBase;
    };

    // MODE 0: strip reductions of the current fields; MODE 3: full tick.
    // KW > 0 (windowed carries): the current reductions are in redHalf[cur];
    // a tick reads them and writes the next ones to redHalf[1 - cur], a
    // MODE 0 pass writes redHalf[cur].  Every item's well has the launch's
    // shape: T threads of CPT columns.
    template <int T, int CPT, int S, int MODE, int KW>
    __global__ __launch_bounds__(T) void BatchStrip(const BatchWell *__restrict__ wells,
                                                    const StripItem *__restrict__ items, float *__restrict__ field,
                                                    int32_t cur)
    {
        const StripItem it = items[blockIdx.x];
        if (it.well < 0)
        {
            return; // padding of the XCD-major work list
        }
        const BatchWell *w = wells + it.well;
        AdiArgs a = w->args;
        a.srcBegin = it.srcBegin;
        a.srcEnd = it.srcEnd;
        if constexpr (KW > 0)
        {
            a.redIn = w->redHalf[cur];
            a.red = (MODE == 0) ? w->redHalf[cur] : w->redHalf[1 - cur];
        }
        if constexpr (CPT == 1)
        {
            StripBody<1, S, MODE, true, KW>(a, field + w->fieldOffset, it.strip);
        }
        else
        {
            if (w->vec)
            {
                StripBody<CPT, S, MODE, true, KW>(a, field + w->fieldOffset, it.strip);
            }
            else
            {
                StripBody<CPT, S, MODE, false, KW>(a, field + w->fieldOffset, it.strip);
            }
        }
    }

    __global__ __launch_bounds__(kScanCols *kScanSegs) void BatchScan(const BatchWell *__restrict__ wells,
                                                                      const ScanItem *__restrict__ items)
    {
        const ScanItem it = items[blockIdx.x];
        const AdiArgs a = wells[it.well].args;
        ScanBody<16>(a, it.yBase, (a.P + kScanSegs - 1) / kScanSegs);
    }

    using BatchStripKernel = void (*)(const BatchWell *, const StripItem *, float *, int32_t);

    template <int T, int CPT, int S>
    BatchStripKernel PickBatchStripRows(int mode, bool window)
    {
        if (window)
        {
            return mode == 0 ? BatchStrip<T, CPT, S, 0, kMaxWindow> : BatchStrip<T, CPT, S, 3, kMaxWindow>;
        }
        return mode == 0 ? BatchStrip<T, CPT, S, 0, 0> : BatchStrip<T, CPT, S, 3, 0>;
    }

    // Strip heights 4 and 8 for every shape; 16 where a thread holds at
    // most 2 columns (the strip's rows stay in registers: rows x columns
    // <= 32 values).
    template <int T, int CPT>
    BatchStripKernel PickBatchStripShape(int rows, int mode, bool window)
    {
        if constexpr (CPT <= 2)
        {
            if (rows == 16)
            {
                return PickBatchStripRows<T, CPT, 16>(mode, window);
            }
        }
        return rows == 4 ? PickBatchStripRows<T, CPT, 4>(mode, window) : PickBatchStripRows<T, CPT, 8>(mode, window);
    }

    bool RowsAllowed(int32_t rows, int32_t cpt)
    {
        return rows == 4 || rows == 8 || (rows == 16 && cpt <= 2);
    }

    BatchStripKernel PickBatchStrip(BlockShape sh, int rows, int mode, bool window)
    {
        switch (sh.threads * 8 + sh.cpt)
        {
            case 64 * 8 + 4:
                return PickBatchStripShape<64, 4>(rows, mode, window);
            case 128 * 8 + 2:
                return PickBatchStripShape<128, 2>(rows, mode, window);
            case 128 * 8 + 4:
                return PickBatchStripShape<128, 4>(rows, mode, window);
            case 256 * 8 + 1:
                return PickBatchStripShape<256, 1>(rows, mode, window);
            case 256 * 8 + 2:
                return PickBatchStripShape<256, 2>(rows, mode, window);
            case 256 * 8 + 4:
                return PickBatchStripShape<256, 4>(rows, mode, window);
            case 512 * 8 + 1:
                return PickBatchStripShape<512, 1>(rows, mode, window);
            case 512 * 8 + 2:
                return PickBatchStripShape<512, 2>(rows, mode, window);
            default:
                return PickBatchStripShape<1024, 1>(rows, mode, window);
        }
    }
}

struct halhip_batch
{
    halhip_context *ctx = nullptr;
    int32_t nWells = 0;
    std::vector<int32_t> X, Y;
    std::vector<int64_t> offset;
    int64_t cells = 0;
    std::vector<double> rate, decay;
    // sources, host copy: (well, cell, value) in list order
    std::vector<int32_t> srcWell;
    std::vector<int64_t> srcCell;
    std::vector<float> srcVal;
    int32_t classRows[kClasses] = {8, 8, 8}; // strip height of each width class
    DeviceArray<float> field;
    DeviceArray<float> tables;
    DeviceArray<float> red;
    DeviceArray<float> carry;
    DeviceArray<float> colInfo;
    DeviceArray<BatchWell> wells;
    DeviceArray<StripItem> stripItems;    // natural order (per-instance ranges)
    DeviceArray<StripItem> stripItemsXcd; // per width class, XCD-major, for batched launches
    int32_t shapeIdx[kClasses] = {0, 0, 0};     // kShapes option of each width class
    int32_t classItemBegin[kClasses] = {0, 0, 0}; // ranges of stripItemsXcd
    int32_t classItemCount[kClasses] = {0, 0, 0};
    std::vector<int32_t> wellClass;
    DeviceArray<ScanItem> scanItems;
    DeviceArray<int32_t> srcXY;
    DeviceArray<float> srcValDev;
    std::vector<int32_t> wellStripBegin, wellStripCount; // per-instance launches: item ranges
    std::vector<int32_t> wellScanBegin, wellScanCount;
    int32_t nStripItems = 0;
    int32_t nScanItems = 0;
    std::string builtKey;
    bool reducedValid = false;
    bool wantWindow = false; // windowed carries requested (halhip_batch_set_carries / autotune)
    bool window = false;     // every well uses windowed carries (no scan launches)
    int32_t cur = 0;     // windowed carries: half of red[] with the current reductions
    StreamClock clock;
    Profiler prof;
    GraphCache graphs;
    bool prepareOnly = false; // inside halhip_batch_prepare_ticks
};

namespace
{
    void RequireBatch(const halhip_batch *b, const char *caller)
    {
        Require(b != nullptr, std::string(caller) + ": null batch");
        Bind(b->ctx);
    }

    BlockShape ShapeOf(const halhip_batch *b, int32_t cls)
    {
        return kShapes[cls][b->shapeIdx[cls]];
    }

    // Builds per-well factor tables, strip constants, work lists and source
    // lists for the current parameters, when they changed.
    void Build(halhip_batch *b, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        Require(bc == nullptr || bc->kind == HALHIP_BC_NONE ||
                    (bc->kind == HALHIP_BC_DIRICHLET && bc->xlo == nullptr && bc->xhi == nullptr && bc->ylo == nullptr &&
                     bc->yhi == nullptr),
                "halhip_batch: only a constant Dirichlet value is supported for batches");
        const bool dir = bc != nullptr && bc->kind == HALHIP_BC_DIRICHLET;
        const double bcValue = dir ? bc->value : 0.0;
        std::string key;
        for (int32_t c = 0; c < kClasses; c++)
        {
            Require(RowsAllowed(b->classRows[c], ShapeOf(b, c).cpt),
                    "halhip_batch: 16-row strips need a workgroup shape of at most 2 columns per thread");
            AppendKey(key, b->classRows[c]);
        }
        AppendKey(key, b->wantWindow);
        for (int32_t c = 0; c < kClasses; c++)
        {
            AppendKey(key, b->shapeIdx[c]);
        }
        AppendKey(key, wrapX);
        AppendKey(key, wrapY);
        AppendKey(key, dir);
        AppendKey(key, bcValue);
        key.append(reinterpret_cast<const char *>(b->rate.data()), b->rate.size() * sizeof(double));
        key.append(reinterpret_cast<const char *>(b->decay.data()), b->decay.size() * sizeof(double));
        AppendKey(key, b->srcVal.size());
        key.append(reinterpret_cast<const char *>(b->srcWell.data()), b->srcWell.size() * sizeof(int32_t));
        key.append(reinterpret_cast<const char *>(b->srcCell.data()), b->srcCell.size() * sizeof(int64_t));
        key.append(reinterpret_cast<const char *>(b->srcVal.data()), b->srcVal.size() * sizeof(float));
        if (key == b->builtKey)
        {
            return;
        }
        Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
        // Sources grouped by well then row (stable), as (x, y) pairs.
        std::vector<int32_t> order(b->srcWell.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t i, int32_t j)
                         {
                             if (b->srcWell[i] != b->srcWell[j])
                             {
                                 return b->srcWell[i] < b->srcWell[j];
                             }
                             const int32_t yi = b->Y[b->srcWell[i]];
                             return b->srcCell[i] / yi < b->srcCell[j] / yi;
                         });
        std::vector<int32_t> xy(2 * order.size());
        std::vector<float> sval(order.size());
        for (size_t k = 0; k < order.size(); k++)
        {
            const int32_t w = b->srcWell[order[k]];
            const int64_t cell = b->srcCell[order[k]];
            xy[2 * k] = static_cast<int32_t>(cell / b->Y[w]);
            xy[2 * k + 1] = static_cast<int32_t>(cell % b->Y[w]);
            sval[k] = b->srcVal[order[k]];
        }
        std::vector<float> host;
        std::vector<BatchWell> wells(static_cast<size_t>(b->nWells));
        std::vector<StripItem> strips;
        std::vector<ScanItem> scans;
        std::vector<size_t> tabOff(static_cast<size_t>(b->nWells));
        struct Offs
        {
            size_t xfw, xib, xbw, xz, yfw, yib, ybw, yz, A, C, E, bxl, bxh, byl, byh;
        };
        std::vector<Offs> offs(static_cast<size_t>(b->nWells));
        std::vector<int64_t> redOff(static_cast<size_t>(b->nWells));
        std::vector<int64_t> colOff(static_cast<size_t>(b->nWells));
        int64_t redTotal = 0;
        int64_t colTotal = 0;
        std::vector<int32_t> windows(static_cast<size_t>(b->nWells), 0);
        size_t srcCursor = 0;
        b->wellStripBegin.assign(b->nWells, 0);
        b->wellStripCount.assign(b->nWells, 0);
        b->wellScanBegin.assign(b->nWells, 0);
        b->wellScanCount.assign(b->nWells, 0);
        b->wellClass.assign(b->nWells, 0);
        for (int32_t w = 0; w < b->nWells; w++)
        {
            const int32_t X = b->X[w];
            const int32_t Yw = b->Y[w];
            const int32_t cls = WidthClass(Yw);
            b->wellClass[w] = cls;
            const BlockShape sh = ShapeOf(b, cls);
            const int32_t S = b->classRows[cls];
            const int32_t cpt = sh.cpt;
            const int32_t P = (X + S - 1) / S;
            const int32_t xPad = P * S;
            const int32_t yPad = sh.threads * cpt;
            AxisFactors fx = FactorLine(X, xPad, b->rate[w], wrapX != 0, dir);
            AxisFactors fy = FactorLine(Yw, yPad, b->rate[w], wrapY != 0, dir);
            auto put = [&](const std::vector<double> &v)
            {
                size_t at = host.size();
                for (double d : v)
                {
                    host.push_back(static_cast<float>(d));
                }
                return at;
            };
            std::vector<double> sA(static_cast<size_t>(P)), sC(static_cast<size_t>(P)), sE(static_cast<size_t>(P));
            for (int32_t q = 0; q < P; q++)
            {
                double A = 1.0;
                double C = 1.0;
                std::vector<double> dp(static_cast<size_t>(S));
                double acc = 1.0;
                for (int32_t r = 0; r < S; r++)
                {
                    const double fwv = fx.fw[static_cast<size_t>(q * S + r)];
                    A *= fwv;
                    C *= fx.bw[static_cast<size_t>(q * S + r)];
                    acc = fwv * acc;
                    dp[r] = acc;
                }
                acc = 0.0;
                for (int32_t r = S - 1; r >= 0; r--)
                {
                    acc = fx.bw[static_cast<size_t>(q * S + r)] * acc + fx.ib[static_cast<size_t>(q * S + r)] * dp[r];
                }
                sA[q] = A;
                sC[q] = C;
                sE[q] = acc;
            }
            windows[w] = wrapX ? 0 : WindowSize(sA, sC);
            Offs &o = offs[w];
            o.xfw = put(fx.fw);
            o.xib = put(fx.ib);
            o.xbw = put(fx.bw);
            o.xz = put(fx.z);
            o.yfw = put(fy.fw);
            o.yib = put(fy.ib);
            o.ybw = put(fy.bw);
            o.yz = put(fy.z);
            o.A = put(sA);
            o.C = put(sC);
            o.E = put(sE);
            o.bxl = put(std::vector<double>(static_cast<size_t>(Yw), bcValue));
            o.bxh = put(std::vector<double>(static_cast<size_t>(Yw), bcValue));
            o.byl = put(std::vector<double>(static_cast<size_t>(X), bcValue));
            o.byh = put(std::vector<double>(static_cast<size_t>(X), bcValue));
            redOff[w] = redTotal;
            redTotal += 2 * static_cast<int64_t>(P) * Yw;
            colOff[w] = colTotal;
            colTotal += 3 * static_cast<int64_t>(Yw);

            BatchWell &bw = wells[w];
            AdiArgs &a = bw.args;
            a = AdiArgs{};
            a.X = X;
            a.Y = Yw;
            a.P = P;
            a.hr = static_cast<float>(b->rate[w] / 2);
            a.r = static_cast<float>(b->rate[w]);
            a.negDecay = static_cast<float>(-b->decay[w]);
            a.wrapX = wrapX ? 1 : 0;
            a.wrapY = wrapY ? 1 : 0;
            a.dirichlet = dir ? 1 : 0;
            a.vecOk = (Yw % cpt == 0) ? 1 : 0;
            a.yvEnd = static_cast<float>(fy.vEnd);
            a.yInv = static_cast<float>(fy.inv);
            a.xvEnd = static_cast<float>(fx.vEnd);
            a.xInv = static_cast<float>(fx.inv);
            a.xibLast = static_cast<float>(fx.ib[static_cast<size_t>(X - 1)]);
            a.xz0 = static_cast<float>(fx.z[0]);
            a.xzLast = static_cast<float>(fx.z[static_cast<size_t>(X - 1)]);
            bw.fieldOffset = b->offset[w];
            bw.vec = a.vecOk;

            b->wellScanBegin[w] = static_cast<int32_t>(scans.size());
            for (int32_t yb = 0; yb < Yw; yb += kScanCols)
            {
                scans.push_back(ScanItem{w, yb});
            }
            b->wellScanCount[w] = static_cast<int32_t>(scans.size()) - b->wellScanBegin[w];
            b->wellStripBegin[w] = static_cast<int32_t>(strips.size());
            for (int32_t q = 0; q < P; q++)
            {
                const int32_t begin = static_cast<int32_t>(srcCursor);
                while (srcCursor < order.size() && b->srcWell[order[srcCursor]] == w &&
                       xy[2 * srcCursor] < (q + 1) * S)
                {
                    srcCursor++;
                }
                strips.push_back(StripItem{w, q, begin, static_cast<int32_t>(srcCursor)});
            }
            b->wellStripCount[w] = static_cast<int32_t>(strips.size()) - b->wellStripBegin[w];
        }
        // Windowed carries when every well allows them: two halves of red[].
        b->window = b->wantWindow && b->nWells > 0 &&
                    std::all_of(windows.begin(), windows.end(), [](int32_t k) { return k > 0; });
        b->cur = 0;
        // Upload tables and fill the device pointers.
        b->tables.Allocate(b->ctx, host.size());
        Check(hipMemcpy(b->tables.get(), host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        b->red.Allocate(b->ctx, static_cast<size_t>(b->window ? 2 * redTotal : redTotal));
        b->carry.Allocate(b->ctx, static_cast<size_t>(redTotal));
        b->colInfo.Allocate(b->ctx, static_cast<size_t>(colTotal));
        Check(hipMemset(b->colInfo.get(), 0, static_cast<size_t>(colTotal) * sizeof(float)), "hipMemset");
        b->srcXY.Allocate(b->ctx, std::max<size_t>(1, xy.size()));
        b->srcValDev.Allocate(b->ctx, std::max<size_t>(1, sval.size()));
        if (!xy.empty())
        {
            Check(hipMemcpy(b->srcXY.get(), xy.data(), xy.size() * sizeof(int32_t), hipMemcpyHostToDevice), "hipMemcpy");
            Check(hipMemcpy(b->srcValDev.get(), sval.data(), sval.size() * sizeof(float), hipMemcpyHostToDevice),
                  "hipMemcpy");
        }
        float *t = b->tables.get();
        for (int32_t w = 0; w < b->nWells; w++)
        {
            AdiArgs &a = wells[w].args;
            const Offs &o = offs[w];
            a.xfw = t + o.xfw;
            a.xib = t + o.xib;
            a.xbw = t + o.xbw;
            a.xz = t + o.xz;
            a.yfw = t + o.yfw;
            a.yib = t + o.yib;
            a.ybw = t + o.ybw;
            a.yz = t + o.yz;
            a.stripA = t + o.A;
            a.stripC = t + o.C;
            a.stripE = t + o.E;
            a.bcXlo = t + o.bxl;
            a.bcXhi = t + o.bxh;
            a.bcYlo = t + o.byl;
            a.bcYhi = t + o.byh;
            a.red = b->red.get() + redOff[w];
            a.carry = b->carry.get() + redOff[w];
            a.window = b->window ? windows[w] : 0;
            wells[w].redHalf[0] = b->red.get() + redOff[w];
            wells[w].redHalf[1] = b->window ? b->red.get() + redTotal + redOff[w] : wells[w].redHalf[0];
            a.colInfo = b->colInfo.get() + colOff[w];
            a.srcXY = b->srcXY.get();
            a.srcVal = b->srcValDev.get();
        }
        b->wells.Allocate(b->ctx, wells.size());
        Check(hipMemcpy(b->wells.get(), wells.data(), wells.size() * sizeof(BatchWell), hipMemcpyHostToDevice),
              "hipMemcpy");
        b->stripItems.Allocate(b->ctx, strips.size());
        Check(hipMemcpy(b->stripItems.get(), strips.data(), strips.size() * sizeof(StripItem), hipMemcpyHostToDevice),
              "hipMemcpy");
        // One work list per width class.  Workgroups are dealt to the 8 XCDs
        // in turn; give XCD x the contiguous run [x*per, (x+1)*per) of a
        // class's list, so that a strip's neighbours (whose reductions the
        // windowed carries read) ran on the same XCD and are likely still in
        // its L2.
        {
            std::vector<StripItem> xcd;
            for (int32_t c = 0; c < kClasses; c++)
            {
                std::vector<StripItem> mine;
                for (const StripItem &it : strips)
                {
                    if (b->wellClass[it.well] == c)
                    {
                        mine.push_back(it);
                    }
                }
                const size_t n = mine.size();
                const size_t per = (n + 7) / 8;
                b->classItemBegin[c] = static_cast<int32_t>(xcd.size());
                b->classItemCount[c] = static_cast<int32_t>(8 * per);
                for (size_t j = 0; j < 8 * per; j++)
                {
                    const size_t src = (j % 8) * per + j / 8;
                    xcd.push_back(src < n ? mine[src] : StripItem{-1, 0, 0, 0});
                }
            }
            b->stripItemsXcd.Allocate(b->ctx, std::max<size_t>(1, xcd.size()));
            if (!xcd.empty())
            {
                Check(hipMemcpy(b->stripItemsXcd.get(), xcd.data(), xcd.size() * sizeof(StripItem),
                                hipMemcpyHostToDevice),
                      "hipMemcpy");
            }
        }
        b->scanItems.Allocate(b->ctx, scans.size());
        Check(hipMemcpy(b->scanItems.get(), scans.data(), scans.size() * sizeof(ScanItem), hipMemcpyHostToDevice),
              "hipMemcpy");
        b->nStripItems = static_cast<int32_t>(strips.size());
        b->nScanItems = static_cast<int32_t>(scans.size());
        b->builtKey = key;
        b->reducedValid = false;
        b->graphs.Clear();
    }

    // Strip kernel of width class cls over `count` items from `items`.
    void LaunchBatchStripItems(halhip_batch *b, int mode, int32_t cls, const StripItem *items, int32_t count,
                               const char *name)
    {
        if (count <= 0)
        {
            return;
        }
        const BlockShape sh = ShapeOf(b, cls);
        b->prof.Begin(b->clock.stream, name);
        hipLaunchKernelGGL(PickBatchStrip(sh, b->classRows[cls], mode, b->window), dim3(static_cast<unsigned>(count)),
                           dim3(static_cast<unsigned>(sh.threads)), 0, b->clock.stream,
                           static_cast<const BatchWell *>(b->wells.get()), items, b->field.get(), b->cur);
        Check(hipGetLastError(), "BatchStrip launch");
        b->prof.End(b->clock.stream);
    }

    // Strip kernels of every width class over their XCD-major work lists.
    void LaunchBatchStripAll(halhip_batch *b, int mode, const char *const (&names)[kClasses])
    {
        for (int32_t c = 0; c < kClasses; c++)
        {
            LaunchBatchStripItems(b, mode, c, b->stripItemsXcd.get() + b->classItemBegin[c], b->classItemCount[c],
                                  names[c]);
        }
    }

    // One well's strips (its range of the natural work list).
    void LaunchBatchStripWell(halhip_batch *b, int mode, int32_t w, const char *name)
    {
        LaunchBatchStripItems(b, mode, b->wellClass[w], b->stripItems.get() + b->wellStripBegin[w],
                              b->wellStripCount[w], name);
    }

    constexpr const char *kTickNames[kClasses] = {"batch_strip_tick_w256", "batch_strip_tick_w512",
                                                  "batch_strip_tick_w1024"};
    constexpr const char *kReduceNames[kClasses] = {"batch_strip_reduce", "batch_strip_reduce",
                                                    "batch_strip_reduce"};

    void LaunchBatchScan(halhip_batch *b, int32_t itemBegin, int32_t itemCount, const char *name)
    {
        b->prof.Begin(b->clock.stream, name);
        hipLaunchKernelGGL(BatchScan, dim3(static_cast<unsigned>(itemCount)), dim3(kScanCols * kScanSegs), 0,
                           b->clock.stream, static_cast<const BatchWell *>(b->wells.get()),
                           static_cast<const ScanItem *>(b->scanItems.get()) + itemBegin);
        Check(hipGetLastError(), "BatchScan launch");
        b->prof.End(b->clock.stream);
    }

    // One tick of every well: the cross-strip scan then the strip kernel, or
    // with windowed carries the strip kernel alone (flipping `cur`).
    void BatchTick(halhip_batch *b)
    {
        if (!b->window)
        {
            LaunchBatchScan(b, 0, b->nScanItems, "batch_scan");
        }
        LaunchBatchStripAll(b, 3, kTickNames);
        if (b->window)
        {
            b->cur = 1 - b->cur;
        }
    }

    void EnqueueBatchTicks(halhip_batch *b, int32_t ticks, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        Require(ticks >= 0, "halhip_batch_run_ticks: negative ticks");
        Build(b, wrapX, wrapY, bc);
        if (ticks == 0)
        {
            return;
        }
        if (!b->reducedValid && !b->prepareOnly)
        {
            LaunchBatchStripAll(b, 0, kReduceNames);
        }
        // 10-tick graph chunks keyed by the reduction half they start from
        // (even: a windowed chunk leaves `cur` as it found it); a prepare
        // call captures both halves and enqueues nothing.
        constexpr int32_t kChunk = 10;
        auto chunkGraph = [&](int32_t half)
        {
            std::string key = b->builtKey;
            AppendKey(key, kChunk);
            AppendKey(key, half);
            const int32_t saved = b->cur;
            b->cur = half;
            hipGraphExec_t exec = FindOrCapture(b->graphs, key, b->clock.stream,
                                                [&]
                                                {
                                                    for (int32_t k = 0; k < kChunk; k++)
                                                    {
                                                        BatchTick(b);
                                                    }
                                                });
            b->cur = saved;
            return exec;
        };
        int32_t done = 0;
        if (!b->prof.On() && ticks >= kChunk)
        {
            if (b->prepareOnly)
            {
                chunkGraph(b->cur);
                if (b->window)
                {
                    chunkGraph(1 - b->cur);
                }
                return;
            }
            hipGraphExec_t exec = chunkGraph(b->cur);
            for (; done + kChunk <= ticks; done += kChunk)
            {
                Check(hipGraphLaunch(exec, b->clock.stream), "hipGraphLaunch");
            }
        }
        if (b->prepareOnly)
        {
            return;
        }
        for (; done < ticks; done++)
        {
            BatchTick(b);
        }
        b->reducedValid = true;
    }

    // Baseline: the same kernels launched well by well (one or two launches
    // per well per tick), the shape of a per-instance loop over single-grid
    // plans.
    void EnqueuePerInstanceTicks(halhip_batch *b, int32_t ticks, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        Require(ticks >= 0, "halhip_batch_run_ticks_per_instance: negative ticks");
        Build(b, wrapX, wrapY, bc);
        if (ticks == 0)
        {
            return;
        }
        for (int32_t w = 0; w < b->nWells; w++)
        {
            if (!b->reducedValid)
            {
                LaunchBatchStripWell(b, 0, w, "instance_strip_reduce");
            }
        }
        for (int32_t t = 0; t < ticks; t++)
        {
            for (int32_t w = 0; w < b->nWells; w++)
            {
                if (!b->window)
                {
                    LaunchBatchScan(b, b->wellScanBegin[w], b->wellScanCount[w], "instance_scan");
                }
                LaunchBatchStripWell(b, 3, w, "instance_strip_tick");
            }
            if (b->window)
            {
                b->cur = 1 - b->cur;
            }
        }
        b->reducedValid = true;
    }

    std::string Signature(const halhip_batch *b)
    {
        uint64_t h = 1469598103934665603ULL;
        auto mix = [&](uint64_t v)
        {
            h ^= v;
            h *= 1099511628211ULL;
        };
        mix(static_cast<uint64_t>(b->nWells));
        for (int32_t w = 0; w < b->nWells; w++)
        {
            mix(static_cast<uint64_t>(b->X[w]) << 32 | static_cast<uint32_t>(b->Y[w]));
        }
        std::ostringstream s;
        s << "gfx950:" << std::hex << h;
        return s.str();
    }
}

extern "C"
{
    int halhip_batch_create(halhip_context *ctx, int32_t nWells, const int32_t *xDims, const int32_t *yDims,
                            halhip_batch **out)
    {
        halhip_batch *b = nullptr;
        int rc = Guard(
            [&]
            {
                Require(out != nullptr, "halhip_batch_create: null out");
                *out = nullptr;
                Require(ctx != nullptr, "halhip_batch_create: null context");
                Require(nWells > 0 && xDims != nullptr && yDims != nullptr, "halhip_batch_create: no wells");
                Bind(ctx);
                b = new halhip_batch();
                b->ctx = ctx;
                b->nWells = nWells;
                b->X.assign(xDims, xDims + nWells);
                b->Y.assign(yDims, yDims + nWells);
                for (int32_t w = 0; w < nWells; w++)
                {
                    Require(b->X[w] > 0 && b->Y[w] > 0, "halhip_batch_create: well dimensions must be positive");
                    if (b->X[w] > kBatchMaxSide || b->Y[w] > kBatchMaxSide)
                    {
                        Fail(HALHIP_ERR_UNSUPPORTED, "halhip_batch_create: well sides are limited to 1024 "
                                                     "(use a single-grid plan for larger grids)");
                    }
                    b->offset.push_back(b->cells);
                    b->cells += static_cast<int64_t>(b->X[w]) * b->Y[w];
                }
                b->rate.assign(static_cast<size_t>(nWells), 0.0);
                b->decay.assign(static_cast<size_t>(nWells), 0.0);
                b->field.Allocate(ctx, static_cast<size_t>(b->cells));
                b->clock.Create();
                Check(hipMemsetAsync(b->field.get(), 0, static_cast<size_t>(b->cells) * sizeof(float), b->clock.stream),
                      "hipMemset");
                Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
                *out = b;
            });
        if (rc != HALHIP_OK && b != nullptr)
        {
            b->clock.Destroy();
            delete b;
        }
        return rc;
    }

    int halhip_batch_destroy(halhip_batch *b)
    {
        return Guard(
            [&]
            {
                if (b == nullptr)
                {
                    return;
                }
                Bind(b->ctx);
                b->graphs.Clear();
                b->clock.Destroy();
                delete b;
            });
    }

    int halhip_batch_set_wells(halhip_batch *b, const double *rates, const double *decays)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_set_wells");
                Require(rates != nullptr, "halhip_batch_set_wells: null rates");
                for (int32_t w = 0; w < b->nWells; w++)
                {
                    Require(std::isfinite(rates[w]) && rates[w] >= 0.0, "halhip_batch_set_wells: bad rate");
                    Require(decays == nullptr || std::isfinite(decays[w]), "halhip_batch_set_wells: bad decay");
                    b->rate[w] = rates[w];
                    b->decay[w] = decays != nullptr ? decays[w] : 0.0;
                }
            });
    }

    int halhip_batch_set_sources(halhip_batch *b, int32_t count, const int32_t *well, const int64_t *cell,
                                 const float *value)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_set_sources");
                Require(count >= 0 && (count == 0 || (well != nullptr && cell != nullptr && value != nullptr)),
                        "halhip_batch_set_sources: bad source arrays");
                for (int32_t k = 0; k < count; k++)
                {
                    Require(well[k] >= 0 && well[k] < b->nWells, "halhip_batch_set_sources: well out of range");
                    Require(cell[k] >= 0 && cell[k] < static_cast<int64_t>(b->X[well[k]]) * b->Y[well[k]],
                            "halhip_batch_set_sources: cell out of range");
                }
                b->srcWell.assign(well, well + count);
                b->srcCell.assign(cell, cell + count);
                b->srcVal.assign(value, value + count);
            });
    }

    int halhip_batch_upload(halhip_batch *b, const float *fields, int64_t n)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_upload");
                Require(fields != nullptr && n == b->cells, "halhip_batch_upload: need the concatenated fields");
                Check(hipMemcpyAsync(b->field.get(), fields, static_cast<size_t>(n) * sizeof(float),
                                     hipMemcpyHostToDevice, b->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
                b->reducedValid = false;
            });
    }

    int halhip_batch_download(halhip_batch *b, float *fields, int64_t n)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_download");
                Require(fields != nullptr && n == b->cells, "halhip_batch_download: need room for the concatenated fields");
                Check(hipMemcpyAsync(fields, b->field.get(), static_cast<size_t>(n) * sizeof(float),
                                     hipMemcpyDeviceToHost, b->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_batch_run_ticks_async(halhip_batch *b, int32_t ticks, int32_t wrapX, int32_t wrapY,
                                     const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_run_ticks");
                b->clock.Mark();
                EnqueueBatchTicks(b, ticks, wrapX, wrapY, bc);
            });
    }

    int halhip_batch_run_ticks_per_instance_async(halhip_batch *b, int32_t ticks, int32_t wrapX, int32_t wrapY,
                                                  const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_run_ticks_per_instance");
                b->clock.Mark();
                EnqueuePerInstanceTicks(b, ticks, wrapX, wrapY, bc);
            });
    }

    int halhip_batch_sync(halhip_batch *b, float *elapsedMs)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_sync");
                b->clock.Sync(elapsedMs);
                b->prof.Collect();
            });
    }

    int halhip_batch_prepare_ticks(halhip_batch *b, int32_t ticks, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_prepare_ticks");
                PrepareScope scope(b->prepareOnly);
                EnqueueBatchTicks(b, ticks, wrapX, wrapY, bc);
            });
    }

    int halhip_batch_run_ticks(halhip_batch *b, int32_t ticks, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        int rc = halhip_batch_run_ticks_async(b, ticks, wrapX, wrapY, bc);
        return rc != HALHIP_OK ? rc : halhip_batch_sync(b, nullptr);
    }

    int halhip_batch_run_ticks_per_instance(halhip_batch *b, int32_t ticks, int32_t wrapX, int32_t wrapY,
                                            const halhip_bc *bc)
    {
        int rc = halhip_batch_run_ticks_per_instance_async(b, ticks, wrapX, wrapY, bc);
        return rc != HALHIP_OK ? rc : halhip_batch_sync(b, nullptr);
    }

    int halhip_batch_set_strip_rows(halhip_batch *b, int32_t rows)
    {
        const int32_t all[kClasses] = {rows, rows, rows};
        return halhip_batch_set_class_rows(b, all);
    }

    int halhip_batch_get_strip_rows(halhip_batch *b, int32_t *rows)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_get_strip_rows");
                Require(rows != nullptr, "halhip_batch_get_strip_rows: null output");
                *rows = b->classRows[kClasses - 1];
            });
    }

    int halhip_batch_set_class_rows(halhip_batch *b, const int32_t *rows)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_set_class_rows");
                Require(rows != nullptr, "halhip_batch_set_class_rows: null rows");
                for (int32_t c = 0; c < kClasses; c++)
                {
                    Require(rows[c] == 4 || rows[c] == 8 || rows[c] == 16,
                            "halhip_batch_set_class_rows: rows must be 4, 8 or 16");
                }
                Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
                for (int32_t c = 0; c < kClasses; c++)
                {
                    if (rows[c] != b->classRows[c])
                    {
                        b->classRows[c] = rows[c];
                        b->builtKey.clear();
                        b->reducedValid = false;
                    }
                }
            });
    }

    int halhip_batch_get_class_rows(halhip_batch *b, int32_t *rows)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_get_class_rows");
                Require(rows != nullptr, "halhip_batch_get_class_rows: null output");
                for (int32_t c = 0; c < kClasses; c++)
                {
                    rows[c] = b->classRows[c];
                }
            });
    }

    int halhip_batch_autotune(halhip_batch *b, int32_t ticks, int32_t wrapX, int32_t wrapY, const halhip_bc *bc,
                              const char *cachePath, int32_t *chosenRows)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_autotune");
                Require(ticks > 0, "halhip_batch_autotune: ticks must be positive");
                const std::string sig = Signature(b);
                std::map<std::string, std::string> cache;
                if (cachePath != nullptr)
                {
                    std::ifstream in(cachePath);
                    std::string k;
                    std::string v;
                    while (in >> k >> v)
                    {
                        cache[k] = v;
                    }
                }
                // A choice: windowed (0/1), then per width class its shape
                // option and strip height, cached as "w:s0.r0,s1.r1,s2.r2".
                struct Choice
                {
                    int32_t window = 0;
                    int32_t shape[kClasses] = {0, 0, 0};
                    int32_t rows[kClasses] = {8, 8, 8};
                };
                auto encode = [](const Choice &c)
                {
                    std::ostringstream o;
                    o << c.window << ":";
                    for (int32_t k = 0; k < kClasses; k++)
                    {
                        o << c.shape[k] << "." << c.rows[k] << (k + 1 < kClasses ? "," : "");
                    }
                    return o.str();
                };
                auto decode = [](const std::string &text, Choice &c)
                {
                    char sep[kClasses * 2 + 1];
                    std::istringstream in(text);
                    in >> c.window >> sep[0];
                    for (int32_t k = 0; k < kClasses; k++)
                    {
                        in >> c.shape[k] >> sep[2 * k + 1] >> c.rows[k];
                        if (k + 1 < kClasses)
                        {
                            in >> sep[2 * k + 2];
                        }
                    }
                    if (!in || (c.window != 0 && c.window != 1))
                    {
                        return false;
                    }
                    for (int32_t k = 0; k < kClasses; k++)
                    {
                        if (c.shape[k] < 0 || c.shape[k] >= kShapesPerClass ||
                            !RowsAllowed(c.rows[k], kShapes[k][c.shape[k]].cpt))
                        {
                            return false;
                        }
                    }
                    return true;
                };
                auto apply = [&](const Choice &c)
                {
                    b->wantWindow = c.window == 1;
                    for (int32_t k = 0; k < kClasses; k++)
                    {
                        b->shapeIdx[k] = c.shape[k];
                        b->classRows[k] = c.rows[k];
                    }
                    b->builtKey.clear();
                    b->reducedValid = false;
                };
                Choice best;
                bool found = false;
                auto it = cache.find(sig);
                if (it != cache.end())
                {
                    found = decode(it->second, best);
                }
                if (!found)
                {
                    // Candidates: carries {scan, windowed} x, per width class,
                    // every (workgroup shape, strip height) its registers
                    // allow.  The width classes run as launches of their own,
                    // so each class's option is timed on its own kernel
                    // events and chosen independently; a carry method costs
                    // the sum of its classes' best times plus its scan
                    // launches.  Candidate j runs option j of every class.
                    // Timing runs on a scratch copy of the fields, so the
                    // batch's state is unchanged.
                    std::vector<std::pair<int32_t, int32_t>> opts[kClasses];
                    size_t nCand = 0;
                    for (int32_t k = 0; k < kClasses; k++)
                    {
                        for (int32_t sh = 0; sh < kShapesPerClass; sh++)
                        {
                            for (int32_t rows : {8, 4, 16})
                            {
                                if (RowsAllowed(rows, kShapes[k][sh].cpt))
                                {
                                    opts[k].push_back({sh, rows});
                                }
                            }
                        }
                        nCand = std::max(nCand, opts[k].size());
                    }
                    DeviceArray<float> saved;
                    saved.Allocate(b->ctx, static_cast<size_t>(b->cells));
                    Check(hipMemcpyAsync(saved.get(), b->field.get(), static_cast<size_t>(b->cells) * sizeof(float),
                                         hipMemcpyDeviceToDevice, b->clock.stream),
                          "hipMemcpyAsync");
                    const bool wasOn = b->prof.On();
                    double bestMs = 0.0;
                    for (int32_t win = 0; win < 2; win++)
                    {
                        Choice cand;
                        cand.window = win;
                        double classMs[kClasses] = {0.0, 0.0, 0.0};
                        bool have[kClasses] = {false, false, false};
                        double scanMs = 0.0;
                        bool anyRun = false;
                        for (size_t j = 0; j < nCand; j++)
                        {
                            Choice c = cand;
                            for (int32_t k = 0; k < kClasses; k++)
                            {
                                const auto &o = opts[k][std::min(j, opts[k].size() - 1)];
                                c.shape[k] = o.first;
                                c.rows[k] = o.second;
                            }
                            apply(c);
                            b->prof.SetOn(false);
                            EnqueueBatchTicks(b, 2, wrapX, wrapY, bc);
                            Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
                            if (win == 1 && !b->window)
                            {
                                // some well's window is too wide at these heights
                                Check(hipMemcpyAsync(b->field.get(), saved.get(),
                                                     static_cast<size_t>(b->cells) * sizeof(float),
                                                     hipMemcpyDeviceToDevice, b->clock.stream),
                                      "hipMemcpyAsync");
                                continue;
                            }
                            b->prof.Reset();
                            b->prof.SetOn(true);
                            EnqueueBatchTicks(b, ticks, wrapX, wrapY, bc);
                            b->prof.Collect();
                            for (int32_t k = 0; k < kClasses; k++)
                            {
                                if (j >= opts[k].size())
                                {
                                    continue; // repeats this class's last option
                                }
                                const double ms = b->prof.TotalMs(kTickNames[k]);
                                if (!have[k] || ms < classMs[k])
                                {
                                    classMs[k] = ms;
                                    have[k] = true;
                                    cand.shape[k] = c.shape[k];
                                    cand.rows[k] = c.rows[k];
                                }
                            }
                            const double sm = b->prof.TotalMs("batch_scan");
                            scanMs = anyRun ? std::min(scanMs, sm) : sm;
                            anyRun = true;
                            Check(hipMemcpyAsync(b->field.get(), saved.get(),
                                                 static_cast<size_t>(b->cells) * sizeof(float),
                                                 hipMemcpyDeviceToDevice, b->clock.stream),
                                  "hipMemcpyAsync");
                        }
                        if (!anyRun || !have[0] || !have[1] || !have[2])
                        {
                            continue;
                        }
                        // the chosen per-class options must also allow the
                        // carry method together
                        apply(cand);
                        EnqueueBatchTicks(b, 1, wrapX, wrapY, bc);
                        Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
                        Check(hipMemcpyAsync(b->field.get(), saved.get(), static_cast<size_t>(b->cells) * sizeof(float),
                                             hipMemcpyDeviceToDevice, b->clock.stream),
                              "hipMemcpyAsync");
                        if (win == 1 && !b->window)
                        {
                            continue;
                        }
                        const double total = scanMs + classMs[0] + classMs[1] + classMs[2];
                        if (!found || total < bestMs)
                        {
                            best = cand;
                            bestMs = total;
                            found = true;
                        }
                    }
                    b->prof.Reset();
                    b->prof.SetOn(wasOn);
                    Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
                    Require(found, "halhip_batch_autotune: no candidate ran");
                    if (cachePath != nullptr)
                    {
                        cache[sig] = encode(best);
                        std::ofstream outFile(cachePath, std::ios::trunc);
                        for (auto &kv : cache)
                        {
                            outFile << kv.first << " " << kv.second << "\n";
                        }
                    }
                }
                apply(best);
                if (chosenRows != nullptr)
                {
                    *chosenRows = b->classRows[kClasses - 1];
                }
            });
    }

    int halhip_batch_set_shapes(halhip_batch *b, const int32_t *shapes)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_set_shapes");
                Require(shapes != nullptr, "halhip_batch_set_shapes: null shapes");
                for (int32_t c = 0; c < kClasses; c++)
                {
                    Require(shapes[c] >= 0 && shapes[c] < kShapesPerClass,
                            "halhip_batch_set_shapes: shape index out of range");
                }
                Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
                for (int32_t c = 0; c < kClasses; c++)
                {
                    b->shapeIdx[c] = shapes[c];
                }
                b->builtKey.clear();
                b->reducedValid = false;
            });
    }

    int halhip_batch_get_shapes(halhip_batch *b, int32_t *shapes, int32_t *threads, int32_t *columnsPerThread)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_get_shapes");
                for (int32_t c = 0; c < kClasses; c++)
                {
                    const BlockShape sh = ShapeOf(b, c);
                    if (shapes != nullptr)
                    {
                        shapes[c] = b->shapeIdx[c];
                    }
                    if (threads != nullptr)
                    {
                        threads[c] = sh.threads;
                    }
                    if (columnsPerThread != nullptr)
                    {
                        columnsPerThread[c] = sh.cpt;
                    }
                }
            });
    }

    int halhip_batch_set_carries(halhip_batch *b, int32_t method)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_set_carries");
                Require(method == HALHIP_ADI_CARRIES_SCAN || method == HALHIP_ADI_CARRIES_WINDOW,
                        "halhip_batch_set_carries: unknown carry method");
                Check(hipStreamSynchronize(b->clock.stream), "hipStreamSynchronize");
                b->wantWindow = method == HALHIP_ADI_CARRIES_WINDOW;
                b->builtKey.clear();
                b->reducedValid = false;
            });
    }

    int halhip_batch_get_carries(halhip_batch *b, int32_t *method)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_get_carries");
                Require(method != nullptr, "halhip_batch_get_carries: null output");
                *method = b->window ? HALHIP_ADI_CARRIES_WINDOW : HALHIP_ADI_CARRIES_SCAN;
            });
    }

    int halhip_batch_set_profiling(halhip_batch *b, int32_t on)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_set_profiling");
                b->prof.Reset();
                b->prof.SetOn(on != 0);
            });
    }

    int halhip_batch_profile(halhip_batch *b, int32_t maxKernels, int32_t *nKernels, char *names, int32_t namesLen,
                             double *totalMs, int64_t *launches)
    {
        return Guard(
            [&]
            {
                RequireBatch(b, "halhip_batch_profile");
                b->prof.Export(maxKernels, nKernels, names, namesLen, totalMs, launches);
            });
    }
}
