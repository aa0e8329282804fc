// sb_driver: the reference's test driver (test/main.cpp:73-158) on this
// library, written against the same public headers and signatures: read two
// OBJ files, prepare both meshes, combine, and write union, difference
// (shifted by (-2, 0, 0)) and intersection (shifted by (2, 0, 0)) into one
// OBJ.  It also prints the face counts and the stage durations.
//
//   sb_driver a.obj b.obj [merged.obj] [--pairs gpu.bin host.bin] [--rays n]
//
// --pairs writes the candidate pairs of the two meshes' trees as the
// reference's search asks for them, test(root, root, boxes, pairs)
// (src/solidboolean.cpp:94-101), once through the GPU search and once through
// the host walk (testNodes), as u64 (a, b) pairs.  --rays n checks the first
// n triangle centroids of the first mesh against the second mesh's tree with
// one-box ray trees as isPointInMesh builds them (src/solidboolean.cpp:55-63),
// against a scan of every box.  Exit status 0 on success, 1 when combine()
// returns false, 2 on a usage or check failure.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "objfile.h"
#include "solidboolean.h"
#include "solidmesh.h"
#include "vector3.h"

namespace {

bool writePairs(const char *path, const std::vector<std::pair<size_t, size_t>> &pairs)
{
    FILE *fp = std::fopen(path, "wb");
    if (!fp)
        return false;
    for (const auto &p : pairs) {
        const unsigned long long ab[2] = {p.first, p.second};
        std::fwrite(ab, sizeof(ab), 1, fp);
    }
    return std::fclose(fp) == 0;
}

long long msBetween(const std::chrono::high_resolution_clock::time_point &a,
                    const std::chrono::high_resolution_clock::time_point &b)
{
    return std::chrono::duration_cast<std::chrono::milliseconds>(b - a).count();
}

}  // namespace

int main(int argc, char **argv)
{
    std::vector<std::string> files;
    const char *gpuPairsPath = nullptr, *hostPairsPath = nullptr;
    long rays = 0;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--pairs") && i + 2 < argc) {
            gpuPairsPath = argv[++i];
            hostPairsPath = argv[++i];
        } else if (!std::strcmp(argv[i], "--rays") && i + 1 < argc) {
            rays = std::atol(argv[++i]);
        } else {
            files.push_back(argv[i]);
        }
    }
    if (files.size() < 2 || files.size() > 3) {
        std::cerr << "usage: sb_driver a.obj b.obj [merged.obj] [--pairs gpu.bin host.bin] [--rays n]" << std::endl;
        return 2;
    }
    std::vector<Vector3> firstVertices, secondVertices;
    std::vector<std::vector<size_t>> firstTriangles, secondTriangles;
    std::string error;
    if (!sbobj::loadObj(files[0], firstVertices, firstTriangles, &error) ||
        !sbobj::loadObj(files[1], secondVertices, secondTriangles, &error)) {
        std::cerr << error << std::endl;
        return 2;
    }

    auto t1 = std::chrono::high_resolution_clock::now();
    SolidMesh firstMesh;
    firstMesh.setVertices(&firstVertices);
    firstMesh.setTriangles(&firstTriangles);
    firstMesh.prepare();
    SolidMesh secondMesh;
    secondMesh.setVertices(&secondVertices);
    secondMesh.setTriangles(&secondTriangles);
    secondMesh.prepare();

    if (gpuPairsPath) {
        const AxisAlignedBoudingBoxTree *left = firstMesh.axisAlignedBoundingBoxTree();
        const AxisAlignedBoudingBoxTree *right = secondMesh.axisAlignedBoundingBoxTree();
        std::vector<std::pair<size_t, size_t>> gpuPairs, hostPairs;
        left->test(left->root(), right->root(), secondMesh.triangleAxisAlignedBoundingBoxes(), &gpuPairs);
        left->testNodes(left->root(), right->root(), secondMesh.triangleAxisAlignedBoundingBoxes(), &hostPairs);
        if (!writePairs(gpuPairsPath, gpuPairs) || !writePairs(hostPairsPath, hostPairs)) {
            std::cerr << "cannot write the pair files" << std::endl;
            return 2;
        }
        std::cout << "tree pairs: " << gpuPairs.size() << " (GPU), " << hostPairs.size() << " (host walk)"
                  << std::endl;
    }
    if (rays > 0) {
        const AxisAlignedBoudingBoxTree *tree = secondMesh.axisAlignedBoundingBoxTree();
        const auto &boxes = *secondMesh.triangleAxisAlignedBoundingBoxes();
        const Vector3 axis(1.7976931348623157e308, 2.220446049250313080847e-16, 2.220446049250313080847e-16);
        size_t checked = 0, candidates = 0;
        for (size_t t = 0; t < firstTriangles.size() && (long)checked < rays; ++t, ++checked) {
            const auto &f = firstTriangles[t];
            const Vector3 p = (firstVertices[f[0]] + firstVertices[f[1]] + firstVertices[f[2]]) / 3.0;
            std::vector<AxisAlignedBoudingBox> rayBox(1);
            rayBox[0].update(p);
            rayBox[0].update(p + axis);
            AxisAlignedBoudingBoxTree testTree(&rayBox, {0}, rayBox[0]);
            std::vector<std::pair<size_t, size_t>> pairs;
            tree->test(tree->root(), testTree.root(), &rayBox, &pairs);
            std::vector<size_t> scan;
            for (size_t i = 0; i < boxes.size(); ++i)
                if (boxes[i].intersectWith(rayBox[0]))
                    scan.push_back(i);
            std::vector<size_t> got;
            for (const auto &pr : pairs)
                got.push_back(pr.first);
            std::sort(got.begin(), got.end());
            if (got != scan) {
                std::cerr << "ray tree of centroid " << t << ": " << got.size() << " candidates, scan "
                          << scan.size() << std::endl;
                return 2;
            }
            candidates += got.size();
        }
        std::cout << "ray trees: " << checked << " rays, " << candidates << " candidates, all equal to a scan"
                  << std::endl;
    }

    SolidBoolean solidBoolean(&firstMesh, &secondMesh);
    const bool ok = solidBoolean.combine();
    auto t2 = std::chrono::high_resolution_clock::now();
    std::cout << "Duration:" << msBetween(t1, t2) << std::endl;
    std::cout << "\tsearchPotentialIntersectedPairs:"
              << msBetween(solidBoolean.benchBegin_searchPotentialIntersectedPairs,
                           solidBoolean.benchEnd_searchPotentialIntersectedPairs)
              << std::endl;
    std::cout << "\tprocessPotentialIntersectedPairs:"
              << msBetween(solidBoolean.benchBegin_processPotentialIntersectedPairs,
                           solidBoolean.benchEnd_processPotentialIntersectedPairs)
              << std::endl;
    std::cout << "\taddUnintersectedTriangles:"
              << msBetween(solidBoolean.benchBegin_addUnintersectedTriangles,
                           solidBoolean.benchEnd_addUnintersectedTriangles)
              << std::endl;
    std::cout << "\treTriangulate:"
              << msBetween(solidBoolean.benchBegin_reTriangulate, solidBoolean.benchEnd_reTriangulate) << std::endl;
    std::cout << "\tbuildPolygonsFromEdges:"
              << msBetween(solidBoolean.benchBegin_buildPolygonsFromEdges,
                           solidBoolean.benchEnd_buildPolygonsFromEdges)
              << std::endl;
    std::cout << "\tbuildFaceGroups:"
              << msBetween(solidBoolean.benchBegin_buildFaceGroups, solidBoolean.benchEnd_buildFaceGroups)
              << std::endl;
    std::cout << "\tdecideGroupSide:"
              << msBetween(solidBoolean.benchBegin_decideGroupSide, solidBoolean.benchEnd_decideGroupSide)
              << std::endl;
    if (!ok) {
        std::cout << "combine failed" << std::endl;
        return 1;
    }

    std::vector<Vector3> mergedVertices;
    std::vector<std::vector<size_t>> mergedTriangles;
    auto mergeMesh = [&](const std::vector<Vector3> &vertices, const std::vector<std::vector<size_t>> &triangles,
                         const Vector3 &offset) {
        const size_t base = mergedVertices.size();
        for (const auto &t : triangles)
            mergedTriangles.push_back({t[0] + base, t[1] + base, t[2] + base});
        for (const auto &v : vertices)
            mergedVertices.push_back(v + offset);
    };
    size_t counts[3];
    {
        std::vector<std::vector<size_t>> resultTriangles;
        solidBoolean.fetchUnion(resultTriangles);
        counts[0] = resultTriangles.size();
        mergeMesh(solidBoolean.resultVertices(), resultTriangles, Vector3(0.0, 0.0, 0.0));
    }
    {
        std::vector<std::vector<size_t>> resultTriangles;
        solidBoolean.fetchDiff(resultTriangles);
        counts[1] = resultTriangles.size();
        mergeMesh(solidBoolean.resultVertices(), resultTriangles, Vector3(-2.0, 0.0, 0.0));
    }
    {
        std::vector<std::vector<size_t>> resultTriangles;
        solidBoolean.fetchIntersect(resultTriangles);
        counts[2] = resultTriangles.size();
        mergeMesh(solidBoolean.resultVertices(), resultTriangles, Vector3(2.0, 0.0, 0.0));
    }
    std::cout << "faces: union " << counts[0] << ", difference " << counts[1] << ", intersection " << counts[2]
              << ", merged " << mergedTriangles.size() << std::endl;
    if (files.size() == 3 && !sbobj::exportObject(files[2], mergedVertices, mergedTriangles)) {
        std::cerr << "cannot write " << files[2] << std::endl;
        return 2;
    }
    return 0;
}
