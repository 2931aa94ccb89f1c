// The graph-motor snapshot's binary face (GMSN), written from a layout's output so a host
// that reads snapshots can take this path's geometry as it takes graph-core's.
//
// Byte contract: docs/contract/binary-layout.md of the reference (graph-contract's
// Snapshot::to_bytes, crates/graph-contract/src/binary.rs:170-197): a 28-byte header
// (magic, format major/minor, node and edge geometry tags, dim, padding, stage count, n, m),
// then node and edge id tables (CSR offsets, UTF-8 bytes, zero padding to 4 bytes), the
// edge endpoints, the node columns x, y (z at 3D), the edge geometry (none for Line) and
// the notes (count, codes, indices). Point nodes and Line edges are what the FA2 layouts
// produce (layout/forceatlas2.rs:82-91); 2D snapshots are labelled 0.3 and 3D ones 0.4
// (snapshot/dim.rs:85-93). Construction refusals follow snapshot.rs: duplicate ids,
// endpoints outside the node table, non-finite coordinates, unsorted or repeated notes.
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/graph_render_fa2.h"

// the calling thread's last-error record (fa2_session.cpp)
extern "C" __attribute__((visibility("hidden"))) gr_status gr_internal_fail(gr_status code,
                                                                             const char *name,
                                                                             const char *rule);

namespace {

void put_u32(std::vector<uint8_t> &b, uint32_t v) {
    for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
}

void put_f32(std::vector<uint8_t> &b, float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    put_u32(b, u);
}

gr_status put_table(std::vector<uint8_t> &b, const char *column, uint32_t count,
                    const char *const *ids) {
    std::unordered_set<std::string> seen;
    std::vector<uint32_t> offs(count + 1, 0);
    std::string text;
    for (uint32_t i = 0; i < count; i++) {
        if (!ids[i]) return gr_internal_fail(GR_ERR_SNAPSHOT, column, "every id is a string");
        std::string s(ids[i]);
        if (!seen.insert(s).second)
            return gr_internal_fail(GR_ERR_SNAPSHOT, column, "every id is unique");
        text += s;
        if (text.size() > 0xFFFFFFFFull)
            return gr_internal_fail(GR_ERR_CAPACITY, column, "id bytes fit a u32 offset");
        offs[i + 1] = (uint32_t)text.size();
    }
    for (uint32_t o : offs) put_u32(b, o);
    b.insert(b.end(), text.begin(), text.end());
    for (size_t p = (4 - text.size() % 4) % 4; p > 0; p--) b.push_back(0);
    return GR_OK;
}

}  // namespace

extern "C" {

gr_status gr_snapshot_encode(uint32_t n, const char *const *node_ids, uint32_t m,
                             const char *const *edge_ids, const uint32_t *source,
                             const uint32_t *target, const float *x, const float *y,
                             const float *z, uint32_t notes, const uint32_t *note_code,
                             const uint32_t *note_index, uint8_t *out, size_t cap,
                             size_t *len_out) {
    if (len_out) *len_out = 0;
    if ((n && (!node_ids || !x || !y)) || (m && (!edge_ids || !source || !target)) ||
        (notes && (!note_code || !note_index)))
        return gr_internal_fail(GR_ERR_PARAM, "snapshot", "every column of a non-empty table");
    std::vector<uint8_t> b;
    b.reserve(28 + (size_t)n * 40 + (size_t)m * 24 + 4 + (size_t)notes * 8);
    b.insert(b.end(), {'G', 'M', 'S', 'N'});
    put_u32(b, 0);           // format major
    put_u32(b, z ? 4u : 3u); // format minor: 0.3 for 2D, 0.4 once a z column is present
    b.push_back(0);          // node geometry: Point
    b.push_back(0);          // edge geometry: Line
    b.push_back(z ? 1 : 0);  // dim
    b.push_back(0);          // padding
    put_u32(b, 1);           // stage count
    put_u32(b, n);
    put_u32(b, m);
    gr_status r;
    if ((r = put_table(b, "node.id", n, node_ids))) return r;
    if ((r = put_table(b, "edge.id", m, edge_ids))) return r;
    for (uint32_t e = 0; e < m; e++)
        if (source[e] >= n || target[e] >= n)
            return gr_internal_fail(GR_ERR_SNAPSHOT,
                                    source[e] >= n ? "edge.source" : "edge.target",
                                    "a position in the node table");
    for (uint32_t e = 0; e < m; e++) put_u32(b, source[e]);
    for (uint32_t e = 0; e < m; e++) put_u32(b, target[e]);
    const float *cols[3] = {x, y, z};
    const char *names[3] = {"node.x", "node.y", "node.z"};
    for (int a = 0; a < (z ? 3 : 2); a++) {
        for (uint32_t i = 0; i < n; i++) {
            if (!std::isfinite(cols[a][i]))
                return gr_internal_fail(GR_ERR_NON_FINITE, names[a], "finite coordinates");
            put_f32(b, cols[a][i]);
        }
    }
    // notes: strictly ascending by (code, index), codes 1-5 (notes.rs:99-176)
    for (uint32_t k = 0; k < notes; k++) {
        const uint32_t c = note_code[k], i = note_index[k];
        if (c < 1 || c > 5) return gr_internal_fail(GR_ERR_SNAPSHOT, "note.code", "codes 1-5");
        if (c == 3 ? i != 0xFFFFFFFFu : i >= m)
            return gr_internal_fail(GR_ERR_SNAPSHOT, "note.index",
                                    "an edge position, or u32::MAX for packing.approximate");
        if (k > 0 && (note_code[k - 1] > c || (note_code[k - 1] == c && note_index[k - 1] >= i)))
            return gr_internal_fail(GR_ERR_SNAPSHOT, "note", "strictly ascending (code, index)");
    }
    put_u32(b, notes);
    for (uint32_t k = 0; k < notes; k++) put_u32(b, note_code[k]);
    for (uint32_t k = 0; k < notes; k++) put_u32(b, note_index[k]);
    if (len_out) *len_out = b.size();
    if (out && cap >= b.size()) std::memcpy(out, b.data(), b.size());
    if (out && cap < b.size())
        return gr_internal_fail(GR_ERR_CAPACITY, "out", "a buffer of *len_out bytes");
    return gr_internal_fail(GR_OK, "", "");
}

}  // extern "C"
