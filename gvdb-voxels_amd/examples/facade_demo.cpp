// facade_demo.cpp -- the point-to-volume sequence of the reference's point
// samples (gPointCloud/gFluidSurface: RebuildTopology, FinishTopology,
// UpdateAtlas, ClearChannel, InsertPointsSubcell, GatherDensity, UpdateApron,
// Render) written against the C++ facade include/gvdb_amd.hpp.
//
// usage: facade_demo <points.f32> <n> <radius> [width]
//   points.f32: n float3 positions, little endian.  Prints one line with the
//   node count per level, the subcell entry count and the mean of the float
//   RGBA image, for comparison with the same run through other bindings, and
//   whether the same stages over two grids through GridBatch give the same
//   entry counts and the same pixels.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gvdb_amd.hpp"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <points.f32> <n> <radius> [width]\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[2]);
    const float radius = float(std::atof(argv[3]));
    const int w = argc > 4 ? std::atoi(argv[4]) : 64;
    std::vector<float> pts(size_t(n) * 3);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(pts.data(), 4, pts.size(), f) != pts.size()) {
        std::fprintf(stderr, "cannot read %d points from %s\n", n, argv[1]);
        return 2;
    }
    std::fclose(f);
    try {
        using namespace gvdb_amd;
        VolumeGVDB gvdb(0);
        gvdb.Configure(5, 4, 3, 3, 3);
        gvdb.SetChannelDefault(16, 16, 1);
        gvdb.AddChannel(0, T_FLOAT, 1);
        gvdb.SetPointsHost(pts.data(), n);
        gvdb.RequestFullRebuild(true);
        gvdb.RebuildTopology(n, radius, Vector3DF(0, 0, 0));
        gvdb.FinishTopology(false, true);
        gvdb.UpdateAtlas();
        gvdb.ClearChannel(0);
        int len = 0;
        gvdb.InsertPointsSubcell(4, n, radius, Vector3DF(0, 0, 0), len);
        gvdb.GatherDensity(4, n, radius, Vector3DF(0, 0, 0), len, 0, 0, false);
        gvdb.UpdateApron(0, 0.0f);
        gvdb.SetCameraOrbit(Vector3DF(30, 20, 0), Vector3DF(256, 256, 256), 600, 1);
        gvdb.SetLight(Vector3DF(556, 756, 456));
        gvdb.SetBackgroundClr(0.1f, 0.1f, 0.15f, 1.0f);
        gvdb.SetCutoff(0.005f, 0.01f, 0.0f);
        gvdb.SetShadowParams(0.8f, 1.0f, 0.0f);
        gvdb.SetExtinct(-0.1f, 1.0f, 0.0f);
        gvdb.SetSteps(0.5f, 16.0f, 0.1f);
        gvdb.SetVolumeRange(0.0f, 0.0f, 2.0f);
        const float a[4] = {1, 1, 1, 0}, b[4] = {1, 0.9f, 0.7f, 1};
        gvdb.LinearTransferFunc(0.0f, 1.0f, a, b);
        gvdb.CommitTransferFunc();
        gvdb.AddRenderBuf(0, w, w, 16);
        gvdb.Render(SHADE_VOLUME, 0, 0);
        std::vector<float> img(size_t(w) * w * 4);
        gvdb.ReadRenderBuf(0, reinterpret_cast<unsigned char*>(img.data()));
        double mean = 0;
        for (float v : img) mean += v;
        mean /= double(img.size());
        std::printf("nodes");
        for (int l = 0; l < 5; l++) std::printf(" %llu", (unsigned long long)gvdb.getNumTotalNodes(l));
        std::printf(" entries %d mean %.9f", len, mean);

        // the same stages over two grids through GridBatch (one launch per kernel)
        VolumeGVDB g1(0), g2(0);
        for (VolumeGVDB* g : {&g1, &g2}) {
            g->Configure(5, 4, 3, 3, 3);
            g->SetChannelDefault(16, 16, 1);
            g->AddChannel(0, T_FLOAT, 1);
            g->SetPointsHost(pts.data(), n);
            g->SetCameraOrbit(Vector3DF(30, 20, 0), Vector3DF(256, 256, 256), 600, 1);
            g->SetLight(Vector3DF(556, 756, 456));
            g->SetBackgroundClr(0.1f, 0.1f, 0.15f, 1.0f);
            g->SetCutoff(0.005f, 0.01f, 0.0f);
            g->SetShadowParams(0.8f, 1.0f, 0.0f);
            g->SetExtinct(-0.1f, 1.0f, 0.0f);
            g->SetSteps(0.5f, 16.0f, 0.1f);
            g->SetVolumeRange(0.0f, 0.0f, 2.0f);
            g->LinearTransferFunc(0.0f, 1.0f, a, b);
            g->CommitTransferFunc();
            g->AddRenderBuf(0, w, w, 16);
        }
        std::vector<VolumeGVDB*> vols = {&g1, &g2};
        GridBatch batch(0, 2);
        batch.RebuildTopology(vols, {n, n}, {radius, radius});
        batch.FinishTopology(vols);
        batch.UpdateAtlas(vols);
        batch.ClearChannel(vols, 0);
        std::vector<int> lens = batch.InsertPointsSubcell(vols, 4, {n, n}, {radius, radius});
        batch.GatherDensity(vols, 4, lens, {radius, radius}, 0);
        batch.UpdateApron(vols, 0, 0.0f);
        batch.Render(vols, SHADE_VOLUME, 0, 0);
        batch.Synchronize();
        std::vector<float> img2(img.size());
        g2.ReadRenderBuf(0, reinterpret_cast<unsigned char*>(img2.data()));
        int same = lens[0] == len && lens[1] == len && img2 == img;
        std::printf(" batched_same %d\n", same);
    } catch (const gvdb_amd::Error& e) {
        std::fprintf(stderr, "gvdb error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
