// oracle_sort.cpp -- TEST INFRASTRUCTURE (CPU baseline timing only).
//
// Stage B2 of the reference exactly as it runs there: std::sort of the
// spatialLookup records, three floats {index, hash, key} each (12 bytes,
// physicsWorld.cc:480), by compareByKey (a.z < b.z, physicsWorld.h:158-161),
// serially (cc:484 passes no execution policy).  The within-bucket order that
// results is libstdc++'s introsort artefact; the parity paths of the oracle
// use the stable (key, index) order instead.
#include <stdint.h>

#include <algorithm>

namespace {
struct Rec {
    float x, y, z;  // index, hash, key as float (glm::vec3)
};
inline bool compare_by_key(const Rec &a, const Rec &b) { return a.z < b.z; }
static_assert(sizeof(Rec) == 12, "12-byte records");
}  // namespace

extern "C" void orc_std_sort_records(float *rec3, uint32_t n)
{
    Rec *r = reinterpret_cast<Rec *>(rec3);
    std::sort(r, r + n, compare_by_key);
}
