/* The bridge runtime behind the bhc C-ABI: queues instructions and flushes
 * them as a BhIR into the first component of the stack (stack level 0).
 * Queueing, the auto-flush threshold and BH_FREE ownership follow the
 * reference bhxx runtime (bridge/cxx/src/Runtime.cpp:28-101). */
#include "../../../include/bhc.h"

#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include <bohrium/bh_component.hpp>
#include <bohrium/bh_main_memory.hpp>

namespace {

using bohrium::component::ComponentFace;

struct Runtime {
    std::unique_ptr<bohrium::ConfigParser> config;
    ComponentFace stack;
    std::vector<bh_instruction> queue;
    std::set<bh_base *> syncs;
    std::vector<bh_base *> freed;  // deleted after the flush that executes their BH_FREE
    std::map<std::string, bh_opcode> extmethods;
    bh_opcode next_ext = BH_MAX_OPCODE_ID + 1;
    int64_t threshold = 1000;
    std::string last_message;

    explicit Runtime(const std::string &fallback) {
        bohrium::ConfigParser::setFallbackPath(fallback);
        config.reset(new bohrium::ConfigParser(-1));
        const std::string top = config->getChildLibraryPath();
        if (top.empty()) throw std::runtime_error("bhc: the configured stack is empty");
        stack = ComponentFace{top, 0};
    }

    // Auto-flush happens before queueing, so the instruction queued last is always pending
    // (bhc_op_slide amends it).  A base enters `freed` only after its BH_FREE is queued, so the
    // auto-flush here never deletes a base whose BH_FREE has not been executed yet (the reference
    // queues before it flushes for the same reason, bridge/cxx/src/Runtime.cpp:33-42).
    void enqueue(bh_instruction &&instr) {
        if (threshold > 0 && static_cast<int64_t>(queue.size()) >= threshold) flush(1, nullptr);
        queue.push_back(std::move(instr));
    }

    void flush(uint64_t nrepeats, bh_base *cond) {
        // Every base in `freed` has its BH_FREE in this batch.
        std::vector<bh_base *> batch_freed;
        batch_freed.swap(freed);
        BhIR bhir(std::move(queue), syncs, nrepeats, cond);
        queue.clear();
        syncs.clear();
        if (recording != nullptr) record(bhir, batch_freed);
        try {
            stack.execute(&bhir);
        } catch (...) {
            release_after_failure(batch_freed);
            throw;
        }
        for (bh_base *b : batch_freed) delete b;
    }

    // ---- recorded flushes (the BhIR wire format, core/bh_ir.cpp) ----
    // File: "BHIRFILE", u32 version 1, then per flush: u64 archive size, the archive, and for each
    // base the archive introduces with host data (in archive order): u64 byte count, the bytes.
    FILE *recording = nullptr;
    std::set<bh_base *> recorded_bases;  // bases the file has introduced and not yet seen freed

    void record(BhIR &bhir, const std::vector<bh_base *> &batch_freed) {
        std::vector<bh_base *> new_data;
        const std::vector<char> a = bhir.writeSerializedArchive(recorded_bases, new_data);
        const uint64_t n = a.size();
        bool ok = std::fwrite(&n, 8, 1, recording) == 1 && std::fwrite(a.data(), 1, a.size(), recording) == a.size();
        for (bh_base *b : new_data) {
            const uint64_t bytes = static_cast<uint64_t>(b->nbytes());
            ok = ok && std::fwrite(&bytes, 8, 1, recording) == 1 &&
                 std::fwrite(b->getDataPtr(), 1, bytes, recording) == bytes;
        }
        ok = ok && std::fflush(recording) == 0;
        // A freed base's address can be reused by a new base, which the file must introduce again.
        for (bh_base *b : batch_freed) recorded_bases.erase(b);
        if (!ok) throw std::runtime_error("bhc: writing the flush recording failed");
    }

    // Bases of replayed flushes, by the address the recording process gave them.
    std::map<const bh_base *, bh_base> replayed;

    void replay(const char *path) {
        FILE *f = std::fopen(path, "rb");
        if (f == nullptr) throw std::runtime_error(std::string("bhc: cannot open ") + path);
        std::unique_ptr<FILE, int (*)(FILE *)> guard(f, std::fclose);
        char magic[8];
        uint32_t version = 0;
        if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "BHIRFILE", 8) != 0 ||
            std::fread(&version, 4, 1, f) != 1 || version != 1)
            throw std::runtime_error(std::string("bhc: not a recorded-flush file: ") + path);
        flush(1, nullptr);  // what this process queued comes first
        uint64_t n = 0;
        while (std::fread(&n, 8, 1, f) == 1) {
            if (n > (uint64_t(1) << 34)) throw std::runtime_error("bhc: recorded flush too large");
            std::vector<char> a(static_cast<size_t>(n));
            if (std::fread(a.data(), 1, a.size(), f) != a.size()) throw std::runtime_error("bhc: truncated recording");
            std::vector<bh_base *> data_recv;
            std::set<bh_base *> frees;
            BhIR bhir(a, replayed, data_recv, frees);
            for (bh_base *b : data_recv) {
                uint64_t bytes = 0;
                if (std::fread(&bytes, 8, 1, f) != 1 || bytes != static_cast<uint64_t>(b->nbytes()))
                    throw std::runtime_error("bhc: recorded data does not match its base");
                void *mem = bh_memory_malloc(bytes);
                if (std::fread(mem, 1, bytes, f) != bytes) {
                    bh_memory_free(mem, bytes);
                    throw std::runtime_error("bhc: truncated recording");
                }
                stack.setMemoryPointer(b, true, mem);
            }
            stack.execute(&bhir);
            for (bh_base *b : frees) replayed.erase(b);  // released by the engine in that flush
        }
    }

    // A failed flush may have stopped before the engine released the bases freed in it.  Hand
    // their BH_FREEs to the stack once more, on their own; a base is deleted only when that
    // succeeds, and leaked otherwise: deleting it unreleased would let a new base at the same
    // address inherit its device buffer.
    void release_after_failure(std::vector<bh_base *> &bases) {
        if (bases.empty()) return;
        std::vector<bh_instruction> frees;
        for (bh_base *b : bases) frees.emplace_back(BH_FREE, std::vector<bh_view>{bh_view(b)});
        BhIR bhir(std::move(frees), std::set<bh_base *>{}, 1, nullptr);
        try {
            stack.execute(&bhir);
        } catch (...) {
            ++leaked_after_failure;
            bases.clear();
            return;
        }
        for (bh_base *b : bases) delete b;
        bases.clear();
    }

    uint64_t leaked_after_failure = 0;
};

Runtime *g_rt = nullptr;
thread_local std::string g_err;

Runtime &rt() {
    if (g_rt == nullptr) throw std::runtime_error("bhc: not initialised (call bhc_init)");
    return *g_rt;
}

bh_view to_view(const bhc_view_t &v) {
    if (v.ndim < 0 || v.ndim > BHC_MAXDIM) throw std::invalid_argument("bhc: view ndim out of range");
    bh_view ret;
    ret.base = static_cast<bh_base *>(v.base);
    ret.start = v.start;
    ret.ndim = v.ndim;
    ret.shape.assign(v.shape, v.shape + v.ndim);
    ret.stride.assign(v.stride, v.stride + v.ndim);
    return ret;
}

// Reject views that reach outside their base: the engine indexes device memory with them.
void check_bounds(const bh_view &v) {
    if (v.isConstant()) return;
    int64_t lo = v.start, hi = v.start;
    bool empty = false;
    for (int64_t d = 0; d < v.ndim; ++d) {
        if (v.shape[d] < 0) throw std::invalid_argument("bhc: negative shape");
        if (v.shape[d] == 0) empty = true;
        const int64_t ext = (v.shape[d] - 1) * v.stride[d];
        if (ext < 0) lo += ext; else hi += ext;
    }
    if (!empty && (lo < 0 || hi >= v.base->nelem()))
        throw std::out_of_range("bhc: view reaches outside its base array");
}

template<typename F>
int guarded(F f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

bh_constant make_constant(int const_type, const void *value) {
    bh_constant c;
    c.type = static_cast<bh_type>(const_type);
    if (const_type < 0 || const_type >= BH_NUM_TYPES) throw std::invalid_argument("bhc: bad constant type");
    std::memcpy(&c.value, value, bh_type_size(c.type));
    return c;
}

} // namespace

extern "C" {

int bhc_init(const char *fallback_config) {
    return guarded([&] {
        if (g_rt == nullptr) g_rt = new Runtime(fallback_config == nullptr ? "" : fallback_config);
    });
}

int bhc_shutdown(void) {
    return guarded([&] {
        if (g_rt == nullptr) return;
        g_rt->flush(1, nullptr);
        if (g_rt->recording != nullptr) std::fclose(g_rt->recording);
        delete g_rt;
        g_rt = nullptr;
    });
}

const char *bhc_last_error(void) { return g_err.c_str(); }

bhc_base_t bhc_new(int dtype, int64_t nelem) {
    bhc_base_t ret = nullptr;
    guarded([&] {
        if (dtype < 0 || dtype >= BH_NUM_TYPES) throw std::invalid_argument("bhc_new: bad dtype");
        if (nelem < 0) throw std::invalid_argument("bhc_new: negative size");
        ret = new bh_base(nelem, static_cast<bh_type>(dtype));
    });
    return ret;
}

int bhc_op(int64_t opcode, int nops, const bhc_view_t *views, int const_type, const void *const_value) {
    return guarded([&] {
        if (!bh_opcode_exists(opcode) || opcode == BH_FREE) throw std::invalid_argument("bhc_op: unknown opcode");
        if (bh_noperands(opcode) != nops) throw std::invalid_argument("bhc_op: wrong number of operands");
        bh_instruction instr;
        instr.opcode = opcode;
        bool has_const = false;
        for (int i = 0; i < nops; ++i) {
            instr.operand.push_back(to_view(views[i]));
            if (instr.operand.back().isConstant()) {
                has_const = true;
            } else {
                check_bounds(instr.operand.back());
            }
        }
        if (has_const) instr.constant = make_constant(const_type, const_value);
        if (nops > 0 && instr.operand[0].isConstant()) throw std::invalid_argument("bhc_op: output is a constant");
        rt().enqueue(std::move(instr));
    });
}

int bhc_random123(const bhc_view_t *out, uint64_t start, uint64_t key) {
    return guarded([&] {
        bh_instruction instr;
        instr.opcode = BH_RANDOM;
        instr.operand.push_back(to_view(*out));
        check_bounds(instr.operand[0]);
        instr.operand.emplace_back();
        instr.constant.type = bh_type::R123;
        instr.constant.value.r123 = {start, key};
        rt().enqueue(std::move(instr));
    });
}

int bhc_op_slide(int operand, int64_t rank, int64_t offset_change, int64_t shape_change, int64_t stride,
                 int64_t shape, int64_t step_delay) {
    return guarded([&] {
        Runtime &r = rt();
        if (r.queue.empty()) throw std::invalid_argument("bhc_op_slide: no queued instruction");
        bh_instruction &i = r.queue.back();
        if (operand < 0 || operand >= static_cast<int>(i.operand.size()) || i.operand[operand].isConstant())
            throw std::invalid_argument("bhc_op_slide: bad operand");
        if (step_delay < 1) throw std::invalid_argument("bhc_op_slide: step_delay must be >= 1");
        bh_slide_dim d;
        d.rank = rank;
        d.offset_change = offset_change;
        d.shape_change = shape_change;
        d.stride = stride;
        d.shape = shape;
        d.step_delay = step_delay;
        i.operand[operand].slides.dims.push_back(d);
    });
}

int bhc_free(bhc_base_t base) {
    return guarded([&] {
        bh_base *b = static_cast<bh_base *>(base);
        bh_instruction instr(BH_FREE, {bh_view(b)});
        Runtime &r = rt();
        for (const auto &kv : r.replayed)
            if (&kv.second == b) throw std::invalid_argument("bhc_free: replayed bases belong to the replay");
        r.syncs.erase(b);
        r.enqueue(std::move(instr));  // may flush the instructions queued before this one
        r.freed.push_back(b);         // deleted after the flush that executes this BH_FREE
    });
}

int bhc_sync(bhc_base_t base) {
    return guarded([&] { rt().syncs.insert(static_cast<bh_base *>(base)); });
}

int bhc_record_flushes(const char *path) {
    return guarded([&] {
        Runtime &r = rt();
        r.flush(1, nullptr);  // the queue so far belongs to no recording
        if (r.recording != nullptr) {
            std::fclose(r.recording);
            r.recording = nullptr;
        }
        r.recorded_bases.clear();
        if (path == nullptr || *path == '\0') return;
        FILE *f = std::fopen(path, "wb");
        if (f == nullptr) throw std::runtime_error(std::string("bhc: cannot write ") + path);
        const uint32_t version = 1;
        if (std::fwrite("BHIRFILE", 1, 8, f) != 8 || std::fwrite(&version, 4, 1, f) != 1) {
            std::fclose(f);
            throw std::runtime_error(std::string("bhc: cannot write ") + path);
        }
        r.recording = f;
    });
}

int bhc_replay_file(const char *path) {
    return guarded([&] { rt().replay(path); });
}

bhc_base_t bhc_replay_base(uint64_t recorded_address) {
    bhc_base_t ret = nullptr;
    guarded([&] {
        auto &m = rt().replayed;
        auto it = m.find(reinterpret_cast<const bh_base *>(static_cast<uintptr_t>(recorded_address)));
        if (it == m.end()) throw std::out_of_range("bhc: no replayed base was recorded at that address");
        ret = &it->second;
    });
    return ret;
}

int bhc_flush(void) {
    return guarded([&] { rt().flush(1, nullptr); });
}

int bhc_flush_and_repeat(uint64_t nrepeats, bhc_base_t condition) {
    return guarded([&] { rt().flush(nrepeats, static_cast<bh_base *>(condition)); });
}

void *bhc_data_get(bhc_base_t base, int copy2host, int force_alloc, int nullify) {
    void *ret = nullptr;
    guarded([&] {
        ret = rt().stack.getMemoryPointer(*static_cast<bh_base *>(base), copy2host != 0, force_alloc != 0, nullify != 0);
    });
    return ret;
}

int bhc_data_set(bhc_base_t base, int host_ptr, void *data) {
    return guarded([&] { rt().stack.setMemoryPointer(static_cast<bh_base *>(base), host_ptr != 0, data); });
}

int bhc_data_copy_in(bhc_base_t base, const void *src, int64_t nbytes) {
    return guarded([&] {
        bh_base *b = static_cast<bh_base *>(base);
        if (nbytes != b->nbytes()) throw std::invalid_argument("bhc_data_copy_in: size mismatch");
        void *mem = bh_memory_malloc(static_cast<uint64_t>(nbytes));
        std::memcpy(mem, src, static_cast<size_t>(nbytes));
        rt().stack.setMemoryPointer(b, true, mem);
    });
}

int bhc_data_copy_out(bhc_base_t base, void *dst, int64_t nbytes) {
    return guarded([&] {
        bh_base *b = static_cast<bh_base *>(base);
        if (nbytes != b->nbytes()) throw std::invalid_argument("bhc_data_copy_out: size mismatch");
        void *p = rt().stack.getMemoryPointer(*b, true, true, false);
        std::memcpy(dst, p, static_cast<size_t>(nbytes));
    });
}

const char *bhc_message(const char *msg) {
    const char *ret = nullptr;
    if (guarded([&] { rt().last_message = rt().stack.message(msg); }) == 0) ret = rt().last_message.c_str();
    return ret;
}

const char *bhc_user_kernel(const char *kernel, int nops, const bhc_view_t *operands, const char *compile_cmd,
                            const char *tag, const char *param) {
    static std::string ret_msg;
    const char *ret = nullptr;
    guarded([&] {
        if (kernel == nullptr || tag == nullptr) throw std::invalid_argument("bhc_user_kernel: null kernel or tag");
        if (nops < 0 || (nops > 0 && operands == nullptr)) throw std::invalid_argument("bhc_user_kernel: bad operands");
        std::vector<bh_view> ops;
        for (int i = 0; i < nops; ++i) {
            ops.push_back(to_view(operands[i]));
            check_bounds(ops.back());
        }
        ret_msg = rt().stack.userKernel(kernel, ops, compile_cmd == nullptr ? "" : compile_cmd, tag,
                                        param == nullptr ? "" : param);
        ret = ret_msg.c_str();
    });
    return ret;
}

int bhc_extmethod(const char *name, int nops, const bhc_view_t *views) {
    return guarded([&] {
        Runtime &r = rt();
        auto it = r.extmethods.find(name);
        bh_opcode opcode;
        if (it == r.extmethods.end()) {
            opcode = r.next_ext++;
            r.stack.extmethod(name, opcode);
            r.extmethods[name] = opcode;
        } else {
            opcode = it->second;
        }
        bh_instruction instr;
        instr.opcode = opcode;
        for (int i = 0; i < nops; ++i) {
            instr.operand.push_back(to_view(views[i]));
            check_bounds(instr.operand.back());
        }
        r.enqueue(std::move(instr));
    });
}

int bhc_set_flush_threshold(int64_t n) {
    return guarded([&] { rt().threshold = n; });
}

int64_t bhc_base_nelem(bhc_base_t base) { return static_cast<bh_base *>(base)->nelem(); }

int bhc_base_dtype(bhc_base_t base) { return static_cast<int>(static_cast<bh_base *>(base)->dtype()); }

} // extern "C"
