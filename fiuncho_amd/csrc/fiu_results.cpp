// Result formats at the boundary: the text line Result<int, float>::str
// writes (include/fiuncho/utils/Result.h:37-45, written one per line by
// app/main.cpp:152-158) and the binary record of Result::serialize /
// deserialize (Result.h:47-66), for callers of the C ABI.  Host only.
#include "../../include/fiuncho_amd.h"

#include <cstdio>
#include <cstring>
#include <string>

#include "fiu_internal.h"

using namespace fiu;

namespace {

// "c_0 c_1 ... c_{k-1} mi": each index then a space, then the value as a
// default-formatted stream writes a float (6 significant digits, %g).
size_t format_line(const int32_t *combo, uint16_t order, float mi, char *buf, size_t cap)
{
    char tmp[32];
    std::string line;
    for (uint16_t q = 0; q < order; ++q) {
        std::snprintf(tmp, sizeof tmp, "%d ", (int)combo[q]);
        line += tmp;
    }
    std::snprintf(tmp, sizeof tmp, "%g", (double)mi);
    line += tmp;
    if (buf && cap) {
        const size_t n = line.size() < cap - 1 ? line.size() : cap - 1;
        std::memcpy(buf, line.data(), n);
        buf[n] = '\0';
    }
    return line.size();
}

constexpr size_t record_bytes(uint16_t order) { return sizeof(size_t) + order * sizeof(int32_t) + sizeof(float); }

}  // namespace

extern "C" {

size_t fiu_result_str(const int32_t *combo, uint16_t order, float mi, char *buf, size_t cap)
{
    if (!combo && order) return 0;
    return format_line(combo, order, mi, buf, cap);
}

size_t fiu_result_serialize(const int32_t *combo, uint16_t order, float mi, void *buf, size_t cap)
{
    const size_t need = record_bytes(order);
    if (!buf || cap < need || (!combo && order)) return need;
    char *p = (char *)buf;
    const size_t k = order;
    std::memcpy(p, &k, sizeof k);
    if (order) std::memcpy(p + sizeof k, combo, order * sizeof(int32_t));
    std::memcpy(p + sizeof k + order * sizeof(int32_t), &mi, sizeof mi);
    return need;
}

fiu_status fiu_result_deserialize(const void *buf, size_t len, int32_t *combo, uint16_t max_order, uint16_t *order,
                                  float *mi, size_t *used)
{
    return guard([&] {
        if (!buf || !order || !mi) fail(FIU_ERR_INVALID, "NULL argument");
        size_t k = 0;
        if (len < sizeof k) fail(FIU_ERR_INVALID, "record shorter than its size field");
        std::memcpy(&k, buf, sizeof k);
        if (k > max_order || k > 0xffff) fail(FIU_ERR_INVALID, "record order above max_order");
        const size_t need = record_bytes((uint16_t)k);
        if (len < need) fail(FIU_ERR_INVALID, "record truncated");
        if (k && !combo) fail(FIU_ERR_INVALID, "combo is NULL");
        const char *p = (const char *)buf;
        if (k) std::memcpy(combo, p + sizeof k, k * sizeof(int32_t));
        std::memcpy(mi, p + sizeof k + k * sizeof(int32_t), sizeof *mi);
        *order = (uint16_t)k;
        if (used) *used = need;
    });
}

fiu_status fiu_write_results(const char *path, const int32_t *combos, const float *mi, uint32_t count,
                             uint16_t order)
{
    return guard([&] {
        if (!path || (count && (!combos || !mi))) fail(FIU_ERR_INVALID, "NULL argument");
        std::FILE *f = std::fopen(path, "w");
        if (!f) fail(FIU_ERR_IO, std::string("cannot open \"") + path + "\" for writing");
        std::string line;
        bool ok = true;
        for (uint32_t i = 0; i < count && ok; ++i) {
            const size_t n = format_line(combos + (size_t)i * order, order, mi[i], nullptr, 0);
            line.resize(n + 1);
            format_line(combos + (size_t)i * order, order, mi[i], &line[0], n + 1);
            line[n] = '\n';
            ok = std::fwrite(line.data(), 1, n + 1, f) == n + 1;
        }
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) fail(FIU_ERR_IO, std::string("write to \"") + path + "\" failed");
    });
}

}  // extern "C"
