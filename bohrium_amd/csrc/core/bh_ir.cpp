/* BhIR wire format: the batch a bridge flushes, as bytes (for shipping a flush to another
 * process, or keeping recorded flushes as fixtures).
 *
 * The reference writes the same content with a Boost binary archive (core/bh_ir.cpp:37-154):
 * repeat count, repeat condition, instruction list, syncs, then the base arrays the receiver
 * has not seen yet, in the order of their first appearance; bases are named by the sender's
 * addresses and the data of new bases travels after the archive.  Boost is not used here, so the
 * layout is this project's own: little-endian, fixed-width fields.
 *
 *   magic "BHIR", u32 version (1)
 *   u64 nrepeats, u64 repeat condition (sender address, 0: none)
 *   u64 instruction count, then per instruction:
 *     i32 opcode, u32 operand count, per operand:
 *       u64 base (sender address, 0: the constant operand), i64 start, i64 ndim,
 *       ndim x i64 shape, ndim x i64 stride,
 *       u64 slide dims, per dim i64 rank, offset_change, shape_change, stride, shape, step_delay,
 *       i64 iteration counter, u64 resets, per reset i64 rank, reset_at, since
 *     i32 constant type, 16 bytes constant value (the bh_constant_value union)
 *   u64 sync count, sync count x u64 base
 *   u64 new-base count, per base: u64 base, i32 type, i64 element count, u8 has data
 */
#include <bohrium/bh_ir.hpp>

#include <cstring>
#include <stdexcept>
#include <string>

namespace {

constexpr uint32_t kVersion = 1;

class Writer {
    std::vector<char> &out;

public:
    explicit Writer(std::vector<char> &o) : out(o) {}
    template <typename T> void put(T v) {
        const char *p = reinterpret_cast<const char *>(&v);
        out.insert(out.end(), p, p + sizeof(T));
    }
    void bytes(const void *p, size_t n) {
        const char *c = static_cast<const char *>(p);
        out.insert(out.end(), c, c + n);
    }
};

class Reader {
    const std::vector<char> &in;
    size_t pos = 0;

public:
    explicit Reader(const std::vector<char> &i) : in(i) {}
    template <typename T> T get() {
        T v;
        bytes(&v, sizeof(T));
        return v;
    }
    void bytes(void *p, size_t n) {
        if (n > in.size() - pos) throw std::runtime_error("BhIR archive: truncated");
        std::memcpy(p, in.data() + pos, n);
        pos += n;
    }
    // A count that must fit in the bytes left (each counted item takes at least `min_item` bytes).
    uint64_t count(size_t min_item) {
        const uint64_t n = get<uint64_t>();
        if (min_item > 0 && n > (in.size() - pos) / min_item) throw std::runtime_error("BhIR archive: bad count");
        return n;
    }
    bool done() const { return pos == in.size(); }
};

uint64_t addr(const bh_base *b) { return static_cast<uint64_t>(reinterpret_cast<uintptr_t>(b)); }
bh_base *remote(uint64_t a) { return reinterpret_cast<bh_base *>(static_cast<uintptr_t>(a)); }

void write_view(Writer &w, const bh_view &v) {
    w.put<uint64_t>(addr(v.base));
    w.put<int64_t>(v.start);
    w.put<int64_t>(v.ndim);
    for (int64_t d = 0; d < v.ndim; ++d) w.put<int64_t>(v.shape[d]);
    for (int64_t d = 0; d < v.ndim; ++d) w.put<int64_t>(v.stride[d]);
    w.put<uint64_t>(v.slides.dims.size());
    for (const bh_slide_dim &s : v.slides.dims) {
        w.put<int64_t>(s.rank);
        w.put<int64_t>(s.offset_change);
        w.put<int64_t>(s.shape_change);
        w.put<int64_t>(s.stride);
        w.put<int64_t>(s.shape);
        w.put<int64_t>(s.step_delay);
    }
    w.put<int64_t>(v.slides.iteration_counter);
    w.put<uint64_t>(v.slides.resets.size());
    for (const auto &r : v.slides.resets) {
        w.put<int64_t>(r.first);
        w.put<int64_t>(r.second.first);
        w.put<int64_t>(r.second.second);
    }
}

bh_view read_view(Reader &r) {
    bh_view v;
    v.base = remote(r.get<uint64_t>());
    v.start = r.get<int64_t>();
    v.ndim = r.get<int64_t>();
    if (v.ndim < 0 || v.ndim > BH_MAXDIM) throw std::runtime_error("BhIR archive: bad view rank");
    v.shape.resize(static_cast<size_t>(v.ndim));
    v.stride.resize(static_cast<size_t>(v.ndim));
    for (int64_t d = 0; d < v.ndim; ++d) v.shape[d] = r.get<int64_t>();
    for (int64_t d = 0; d < v.ndim; ++d) v.stride[d] = r.get<int64_t>();
    const uint64_t ndims = r.count(48);
    for (uint64_t i = 0; i < ndims; ++i) {
        bh_slide_dim s;
        s.rank = r.get<int64_t>();
        s.offset_change = r.get<int64_t>();
        s.shape_change = r.get<int64_t>();
        s.stride = r.get<int64_t>();
        s.shape = r.get<int64_t>();
        s.step_delay = r.get<int64_t>();
        v.slides.dims.push_back(s);
    }
    v.slides.iteration_counter = r.get<int64_t>();
    const uint64_t nresets = r.count(24);
    for (uint64_t i = 0; i < nresets; ++i) {
        const int64_t rank = r.get<int64_t>();
        const int64_t at = r.get<int64_t>();
        const int64_t since = r.get<int64_t>();
        v.slides.resets[rank] = {at, since};
    }
    return v;
}

} // namespace

std::vector<char> BhIR::writeSerializedArchive(std::set<bh_base *> &known_base_arrays, std::vector<bh_base *> &new_data) {
    std::vector<bh_base *> new_bases;
    for (bh_instruction &instr : instr_list) {
        for (const bh_view *v : instr.getViews()) {
            if (known_base_arrays.insert(v->base).second) {
                new_bases.push_back(v->base);
                if (v->base->getDataPtr() != nullptr) new_data.push_back(v->base);
            }
        }
    }
    std::vector<char> ret;
    Writer w(ret);
    w.bytes("BHIR", 4);
    w.put<uint32_t>(kVersion);
    w.put<uint64_t>(_nrepeats);
    const bool cond_known = _repeat_condition != nullptr && known_base_arrays.count(_repeat_condition) > 0;
    w.put<uint64_t>(cond_known ? addr(_repeat_condition) : 0);
    w.put<uint64_t>(instr_list.size());
    for (const bh_instruction &instr : instr_list) {
        w.put<int32_t>(static_cast<int32_t>(instr.opcode));
        w.put<uint32_t>(static_cast<uint32_t>(instr.operand.size()));
        for (const bh_view &v : instr.operand) write_view(w, v);
        w.put<int32_t>(static_cast<int32_t>(instr.constant.type));
        static_assert(sizeof(bh_constant_value) == 16, "constant value is 16 bytes");
        w.bytes(&instr.constant.value, sizeof(bh_constant_value));
    }
    w.put<uint64_t>(_syncs.size());
    for (bh_base *b : _syncs) w.put<uint64_t>(addr(b));
    w.put<uint64_t>(new_bases.size());
    for (bh_base *b : new_bases) {
        w.put<uint64_t>(addr(b));
        w.put<int32_t>(static_cast<int32_t>(b->dtype()));
        w.put<int64_t>(b->nelem());
        w.put<uint8_t>(b->getDataPtr() != nullptr ? 1 : 0);
    }
    return ret;
}

BhIR::BhIR(const std::vector<char> &serialized_archive, std::map<const bh_base *, bh_base> &remote2local,
           std::vector<bh_base *> &data_recv, std::set<bh_base *> &frees)
    : _nrepeats(1), _repeat_condition(nullptr) {
    Reader r(serialized_archive);
    char magic[4];
    r.bytes(magic, 4);
    if (std::memcmp(magic, "BHIR", 4) != 0) throw std::runtime_error("BhIR archive: bad magic");
    const uint32_t version = r.get<uint32_t>();
    if (version != kVersion) throw std::runtime_error("BhIR archive: unsupported version " + std::to_string(version));
    _nrepeats = r.get<uint64_t>();
    _repeat_condition = remote(r.get<uint64_t>());
    const uint64_t ninstr = r.count(8);
    instr_list.reserve(static_cast<size_t>(ninstr));
    for (uint64_t i = 0; i < ninstr; ++i) {
        bh_instruction instr;
        instr.opcode = static_cast<bh_opcode>(r.get<int32_t>());
        const uint32_t nops = r.get<uint32_t>();
        if (nops > 64) throw std::runtime_error("BhIR archive: bad operand count");
        for (uint32_t k = 0; k < nops; ++k) instr.operand.push_back(read_view(r));
        const int32_t ctype = r.get<int32_t>();
        if (ctype < 0 || ctype > static_cast<int32_t>(bh_type::R123))
            throw std::runtime_error("BhIR archive: bad constant type");
        instr.constant.type = static_cast<bh_type>(ctype);
        r.bytes(&instr.constant.value, sizeof(bh_constant_value));
        instr_list.push_back(std::move(instr));
    }
    const uint64_t nsyncs = r.count(8);
    std::vector<bh_base *> syncs;
    for (uint64_t i = 0; i < nsyncs; ++i) syncs.push_back(remote(r.get<uint64_t>()));
    const uint64_t nnew = r.count(21);
    struct New {
        uint64_t addr;
        bh_base base;
        bool data;
    };
    std::vector<New> news;
    for (uint64_t i = 0; i < nnew; ++i) {
        New n;
        n.addr = r.get<uint64_t>();
        const int32_t t = r.get<int32_t>();
        const int64_t nelem = r.get<int64_t>();
        n.data = r.get<uint8_t>() != 0;
        if (t < 0 || t >= static_cast<int32_t>(bh_type::R123) || nelem < 0)
            throw std::runtime_error("BhIR archive: bad base");
        n.base = bh_base(nelem, static_cast<bh_type>(t));
        news.push_back(n);
    }
    if (!r.done()) throw std::runtime_error("BhIR archive: trailing bytes");

    for (const bh_instruction &instr : instr_list)
        if (instr.opcode == BH_FREE && !instr.operand.empty()) frees.insert(instr.operand[0].base);
    // The new bases, in the order the sender met them.
    size_t next = 0;
    for (const bh_instruction &instr : instr_list) {
        for (const bh_view *v : instr.getViews()) {
            if (remote2local.count(v->base)) continue;
            if (next >= news.size() || news[next].addr != addr(v->base))
                throw std::runtime_error("BhIR archive: a base is used before it is introduced");
            bh_base &local = remote2local.emplace(v->base, news[next].base).first->second;
            if (news[next].data) data_recv.push_back(&local);
            ++next;
        }
    }
    if (next != news.size()) throw std::runtime_error("BhIR archive: unused new bases");
    for (bh_instruction &instr : instr_list)
        for (bh_view *v : instr.getViews()) v->base = &remote2local.at(v->base);
    for (bh_base *b : syncs) {
        auto it = remote2local.find(b);
        if (it != remote2local.end()) _syncs.insert(&it->second);
    }
    if (_repeat_condition != nullptr) {
        auto it = remote2local.find(_repeat_condition);
        if (it == remote2local.end()) throw std::runtime_error("BhIR archive: unknown repeat condition");
        _repeat_condition = &it->second;
    }
}
