// Host side of the Huffman coder: reference-tree codebooks from GPU histograms,
// HUF<T> header serialization, and multi-level decode tables for the GPU
// decoder.  Stream format: huffman.hpp:322-381 (sizes, payload), 650-783 (tree).
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "internal.hpp"

namespace asz {

namespace {

void put_be64(std::vector<uint8_t>& o, uint64_t v) {
    for (int k = 7; k >= 0; k--) o.push_back((uint8_t)(v >> (8 * k)));
}
void put_le64(std::vector<uint8_t>& o, uint64_t v) {
    for (int k = 0; k < 8; k++) o.push_back((uint8_t)(v >> (8 * k)));
}
uint64_t get_be64(const uint8_t* p) {
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v = (v << 8) | p[k];
    return v;
}
uint64_t get_le64(const uint8_t* p) {
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v |= (uint64_t)p[k] << (8 * k);
    return v;
}

// LSB-first bit writer (n <= 32 bits per call).
struct BitSink {
    std::vector<uint8_t>& o;
    uint64_t acc = 0;
    int fill = 0;
    void put(uint64_t v, int n) {
        acc |= (v & ((1ull << n) - 1)) << fill;
        fill += n;
        while (fill >= 8) { o.push_back((uint8_t)acc); acc >>= 8; fill -= 8; }
    }
    void flush() { if (fill) { o.push_back((uint8_t)acc); acc = 0; fill = 0; } }
};

// Ascending sort of 64-bit keys, LSD radix with 11-bit digits over the bits in use.
void radix_sort(std::vector<uint64_t>& k) {
    if (k.size() < 64) { std::sort(k.begin(), k.end()); return; }
    uint64_t all = 0;
    for (uint64_t v : k) all |= v;
    std::vector<uint64_t> tmp(k.size());
    for (int sh = 0; sh < 64 && (all >> sh); sh += 11) {
        size_t cnt[2049] = {0};
        for (uint64_t v : k) cnt[((v >> sh) & 2047) + 1]++;
        for (int d = 0; d < 2048; d++) cnt[d + 1] += cnt[d];
        for (uint64_t v : k) tmp[cnt[(v >> sh) & 2047]++] = v;
        k.swap(tmp);
    }
}

constexpr uint64_t kLenMask = 0x1234abcdull;

// The header's usemp bit and maxval field follow T's arithmetic
// (huffman.hpp:275-281, 660): T = int32_t for QB arrays, int64_t for t.
template <typename T>
void type_fields(int64_t offset, uint64_t range, uint64_t count, bool& usemp, uint64_t& maxval_field, int64_t& mx) {
    using U = typename std::make_unsigned<T>::type;
    T mn = (T)offset;
    T tmx = (T)((U)mn + (U)(range - 1));
    T maxval = (T)((U)tmx - (U)mn + (U)1);
    usemp = ((int64_t)maxval >= (1 << 12) && count < 2 * (uint64_t)(int64_t)tmx) || (int64_t)maxval >= (1 << 28);
    maxval_field = (uint64_t)(int64_t)maxval;
    mx = (int64_t)tmx;
}

}  // namespace

void pack_book(HuffArray& a) {
    a.book.resize(a.range);
    a.code32.resize(a.range);
    int ml = 0;
    for (uint64_t s = 0; s < a.range; s++) {
        if (a.len[s] > 56) fail(ASZ_ERR_DATA, "Huffman code longer than 56 bits");
        a.book[s] = a.code[s] | ((uint64_t)a.len[s] << 56);
        a.code32[s] = (uint32_t)a.code[s];
        ml = std::max<int>(ml, a.len[s]);
    }
    a.max_len = ml;
}

void build_codebook(HuffArray& a, const uint32_t* hist) {
    a.header.clear();
    a.code.assign(a.range, 0);
    a.len.assign(a.range, 0);
    a.payload_bits = 0;
    bool usemp;
    uint64_t maxval_field;
    int64_t mx;
    if (a.is_int64) type_fields<int64_t>(a.offset, a.range, a.count, usemp, maxval_field, mx);
    else type_fields<int32_t>(a.offset, a.range, a.count, usemp, maxval_field, mx);

    // leaves in ascending symbol order (huffman.hpp:284-303)
    std::vector<uint32_t> sym;
    std::vector<uint64_t> freq;
    for (uint64_t s = 0; s < a.range; s++)
        if (hist[s]) { sym.push_back((uint32_t)s); freq.push_back(hist[s]); }
    const uint64_t n = sym.size();
    if (n == 0) fail(ASZ_ERR_DATA, "empty Huffman histogram");
    if (n == 1 || maxval_field == 1) {  // huffman.hpp:193-210, 342-347
        a.single = true;
        a.header.push_back((uint8_t)((usemp << 7) | (1 << 6) | 1));
        put_le64(a.header, (uint64_t)(a.offset + (int64_t)sym[0]));
        put_be64(a.header, n);
        put_be64(a.header, 1);
        put_be64(a.header, a.count ^ kLenMask);
        a.len[sym[0]] = 0;
        return;
    }
    a.single = false;
    int mbft = 1;
    while ((1ull << mbft) < maxval_field) ++mbft;

    // Reference merge order: repeatedly join the two smallest (freq, node
    // index) nodes, first popped = child 0 (huffman.hpp:216-235).  Leaves come
    // sorted by (freq, index); joined nodes are created with non-decreasing
    // freq and increasing index, so two FIFO queues give the same sequence.
    std::vector<uint32_t> lq(n);
    {
        std::vector<uint64_t> key(n);  // (freq, leaf index), ascending
        for (uint64_t i = 0; i < n; i++) key[i] = (freq[i] << 32) | i;
        radix_sort(key);
        for (uint64_t i = 0; i < n; i++) lq[i] = (uint32_t)key[i];
    }
    const uint64_t total_nodes = 2 * n - 1;
    std::vector<uint64_t> nf(total_nodes);
    std::vector<int64_t> left(total_nodes, -1), right(total_nodes, -1);
    for (uint64_t i = 0; i < n; i++) nf[i] = freq[i];
    uint64_t qi = 0, ii = n, next = n;  // leaf queue head, internal queue head, next internal index
    auto pop = [&]() -> uint64_t {
        bool take_leaf;
        if (qi >= n) take_leaf = false;
        else if (ii >= next) take_leaf = true;
        else take_leaf = nf[lq[qi]] <= nf[ii];  // equal freq: leaf index < internal index
        return take_leaf ? lq[qi++] : ii++;
    };
    while (next < total_nodes) {
        uint64_t u = pop();
        uint64_t v = pop();
        nf[next] = nf[u] + nf[v];
        left[next] = (int64_t)u;
        right[next] = (int64_t)v;
        next++;
    }
    const uint64_t root = total_nodes - 1;

    // codes: bit j of the code is the branch taken at depth j (huffman.hpp:80-95)
    {
        struct Item { uint64_t node, code; int len; };
        std::vector<Item> st{{root, 0, 0}};
        while (!st.empty()) {
            Item it = st.back();
            st.pop_back();
            if (left[it.node] < 0) {
                a.code[sym[it.node]] = it.code;
                a.len[sym[it.node]] = (uint8_t)it.len;
                a.payload_bits += freq[it.node] * (uint64_t)it.len;
                continue;
            }
            if (it.len >= 64) fail(ASZ_ERR_DATA, "Huffman code longer than 64 bits");
            st.push_back({(uint64_t)right[it.node], it.code | (1ull << it.len), it.len + 1});
            st.push_back({(uint64_t)left[it.node], it.code, it.len + 1});
        }
    }
    // header + DFS preorder tree (huffman.hpp:654-690)
    a.header.push_back((uint8_t)((usemp << 7) | mbft));
    put_le64(a.header, (uint64_t)a.offset);
    put_be64(a.header, n);
    put_be64(a.header, maxval_field);
    BitSink bs{a.header};
    std::vector<uint64_t> st{root};
    while (!st.empty()) {
        uint64_t u = st.back();
        st.pop_back();
        if (left[u] < 0) { bs.put(1, 1); bs.put(sym[u], mbft); }
        else { bs.put(0, 1); st.push_back((uint64_t)right[u]); st.push_back((uint64_t)left[u]); }
    }
    bs.flush();
    put_be64(a.header, a.payload_bits ^ kLenMask);
}

// ------------------------------------------------------------------ decoding tables
//
// Multi-level lookup tables over the LSB-first bit stream.  Entry (uint32):
//   leaf:      bit 31 = 0, bits 0..23 = symbol index (sym - offset) if < 2^24,
//              bits 24..30 = code length consumed at this level (1..64)
//   subtable:  bit 31 = 1, bits 0..23 = offset of the subtable in words,
//              bits 24..28 = subtable index bits
// Symbol indices >= 2^24 use a side array: leaf entries then hold
// 0xFFFFFF and the symbol index lives at table[sub_base + ...] -- not needed
// for ranges below 2^24, which every stream the core writes satisfies; larger
// alphabets are rejected with an error.

namespace {

// One leaf of a serialized tree: symbol index, code (LSB-first: the branch
// taken at depth k is bit k, 1 = right) and length.
struct Leaf {
    uint64_t code;
    uint32_t sym;
    uint32_t len;
};

// Fill the table of 2^bits entries at tab[base] for leaves lv[a, b), all of
// which share their first `used` code bits: a leaf with at most used + bits
// bits fills every entry whose low (len - used) bits are its remaining code;
// longer leaves, contiguous per (used + bits)-bit prefix because the leaves
// are in preorder, get a subtable of min(max length - used - bits,
// kSubTableBits) bits reached through an entry with the top bit set.
void fill_level(const std::vector<Leaf>& lv, uint64_t a, uint64_t b, int used, int bits, std::vector<uint32_t>& tab,
                uint64_t base) {
    const uint64_t mask = (1ull << bits) - 1;
    for (uint64_t i = a; i < b;) {
        const Leaf& L = lv[i];
        const int rem = (int)L.len - used;
        const uint64_t code = L.code >> used;
        if (rem <= bits) {
            const uint32_t e = L.sym | ((uint32_t)rem << 24);
            for (uint64_t q = code & mask; q <= mask; q += 1ull << rem) tab[base + q] = e;
            i++;
            continue;
        }
        uint64_t j = i + 1;
        uint32_t mx = L.len;
        while (j < b && (int)lv[j].len - used > bits && ((lv[j].code >> used) & mask) == (code & mask))
            mx = std::max(mx, lv[j++].len);
        const int sb = std::min((int)mx - used - bits, kSubTableBits);
        const uint64_t off = tab.size();
        if (off + (1ull << sb) > 0xFFFFFFull) fail(ASZ_ERR_DATA, "Huffman decode table too large");
        tab.resize(off + (1ull << sb), 0);
        tab[base + (code & mask)] = 0x80000000u | (uint32_t)off | ((uint32_t)sb << 24);
        fill_level(lv, i, j, used + bits, sb, tab, off);
        i = j;
    }
}

}  // namespace

DecodeInfo parse_huffman(const uint8_t* s, uint64_t n, uint64_t off, uint64_t count) {
    DecodeInfo di;
    if (off + 25 > n) fail(ASZ_ERR_DATA, "compressed stream is truncated in a Huffman header");
    const uint8_t* p = s + off;
    int mbft = p[0] & 0x3f;
    di.offset = (int64_t)get_le64(p + 1);
    di.nleaf = get_be64(p + 9);
    uint64_t maxval = get_be64(p + 17);
    uint64_t pos = off + 25;
    if (di.nleaf == 0) fail(ASZ_ERR_DATA, "Huffman tree has no symbols");
    if (maxval == 1 || di.nleaf == 1) {
        if (pos + 8 > n) fail(ASZ_ERR_DATA, "compressed stream is truncated in a Huffman payload");
        di.single = true;
        di.single_count = get_be64(s + pos) ^ kLenMask;
        if (count && di.single_count != count)
            fail(ASZ_ERR_DATA, "Huffman payload length does not match its value count");
        di.end_off = pos + 8;
        return di;
    }
    if (mbft > 31) fail(ASZ_ERR_DATA, "Huffman symbol width too large");
    // walk the DFS preorder tree (huffman.hpp:741-774) without building it:
    // bit 0 is the root marker (the root is internal), then per node 0 =
    // internal, 1 = leaf followed by its mbft-bit symbol.  The position is
    // the path of d branches, kept MSB-first in m and LSB-first (the code) in
    // c: an internal node descends left (m <<= 1, c unchanged); after a leaf
    // the walk climbs the t trailing right branches and steps to the right
    // sibling (m = (m >> t) + 1; c keeps its low d - t - 1 bits and sets bit
    // d - t - 1), and ends when it climbs past the root.
    const uint64_t avail_bits = (n - pos) * 8;
    auto need = [&](uint64_t k) {
        if (k > avail_bits) fail(ASZ_ERR_DATA, "compressed stream is truncated in a Huffman tree");
    };
    // m <= 32 bits from bit k on, LSB-first: up to five bytes in one word
    auto bits = [&](uint64_t k, int m) -> uint64_t {
        need(k + m);
        const uint64_t b0 = pos + (k >> 3), nb = ((k & 7) + m + 7) >> 3;
        uint64_t w = 0;
        for (uint64_t j = 0; j < nb; j++) w |= (uint64_t)s[b0 + j] << (8 * j);
        return (w >> (k & 7)) & ((1ull << m) - 1);
    };
    std::vector<Leaf> lv;
    lv.reserve(di.nleaf);
    uint64_t i = 1, m = 0, c = 0;
    int d = 1;
    uint32_t max_len = 0;
    for (;;) {
        // a run of internal nodes, then a leaf
        const int chunk = (int)std::min<uint64_t>(32, avail_bits > i ? avail_bits - i : 0);
        const uint64_t w = chunk ? bits(i, chunk) : 0;
        if (w == 0) {
            need(i + chunk + 1);  // no leaf in the bits left: truncated
            d += chunk;
            m <<= std::min(chunk, 63);
            i += chunk;
            if (d > 64) fail(ASZ_ERR_DATA, "Huffman code longer than 64 bits");
            continue;
        }
        const int z = __builtin_ctzll(w);
        d += z;
        if (d > 64) fail(ASZ_ERR_DATA, "Huffman code longer than 64 bits");
        m = z ? m << z : m;
        i += z + 1;
        const uint64_t sym = mbft ? bits(i, mbft) : 0;
        i += mbft;
        if (sym >= (1ull << 24)) fail(ASZ_ERR_DATA, "Huffman alphabet too wide for the GPU decoder");
        if (lv.size() == di.nleaf) fail(ASZ_ERR_DATA, "malformed Huffman tree");
        lv.push_back({c, (uint32_t)sym, (uint32_t)d});
        max_len = std::max(max_len, (uint32_t)d);
        const int t = (~m == 0) ? 64 : __builtin_ctzll(~m);
        if (t >= d) break;
        m = (m >> t) + 1;
        d -= t;
        c = (c & ((1ull << (d - 1)) - 1)) | (1ull << (d - 1));
    }
    const uint64_t leaves = lv.size();
    if (leaves != di.nleaf) fail(ASZ_ERR_DATA, "Huffman tree leaf count mismatch");
    pos += (i + 7) >> 3;
    if (pos + 8 > n) fail(ASZ_ERR_DATA, "compressed stream is truncated in a Huffman payload");
    di.payload_bits = get_be64(s + pos) ^ kLenMask;
    pos += 8;
    di.payload_off = pos;
    di.end_off = pos + ((di.payload_bits + 7) >> 3);
    if (di.end_off > n) fail(ASZ_ERR_DATA, "compressed stream is truncated in a Huffman payload");
    di.max_len = (int)max_len;
    const int rb = std::min(di.max_len, kRootBits);
    di.table.assign(1ull << rb, 0);
    fill_level(lv, 0, leaves, 0, rb, di.table, 0);
    return di;
}

}  // namespace asz

extern "C" int asz_huffman_codebook(const uint64_t* hist, uint64_t range, int64_t offset, uint64_t count,
                                    int is_int64, uint8_t* hdr, uint64_t hdr_cap, uint64_t* hdr_bytes,
                                    uint64_t* codes, uint8_t* lens, uint64_t* payload_bits) {
    try {
        asz::HuffArray a;
        a.count = count;
        a.offset = offset;
        a.range = range;
        a.is_int64 = is_int64 != 0;
        std::vector<uint32_t> h32(range);
        for (uint64_t i = 0; i < range; i++) {
            if (hist[i] > 0xFFFFFFFFull) asz::fail(ASZ_ERR_ARG, "histogram count exceeds 32 bits");
            h32[i] = (uint32_t)hist[i];
        }
        asz::build_codebook(a, h32.data());
        if (a.header.size() > hdr_cap) asz::fail(ASZ_ERR_ARG, "header buffer too small");
        std::memcpy(hdr, a.header.data(), a.header.size());
        *hdr_bytes = a.header.size();
        for (uint64_t i = 0; i < range; i++) { codes[i] = a.code[i]; lens[i] = a.len[i]; }
        *payload_bits = a.single ? 0 : a.payload_bits;
        return ASZ_OK;
    } catch (const asz::Error& e) {
        return e.code;
    } catch (...) {
        return ASZ_ERR_DATA;
    }
}
