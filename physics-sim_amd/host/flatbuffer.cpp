// flatbuffer.cpp - reader and writer of the reference's saved-state flatbuffer.
//
// Table `State` (new_schema.fbs:1-105): 16 fields in slot order pos, vel, acc, rot,
// rot_vel, rot_acc, forces, radii, fixity, bonds, contacts, material_pointers,
// materials, groups, scripts, settings (voffset 4 + 2*slot, new_schema_generated.rs
// :2500-2515); `settings` is an inline 212-byte struct.  The older schema.fbs has 15
// slots, no `scripts` and a 204-byte struct in the last slot.
//
// The writer follows the published FlatBuffers back-to-front algorithm (alignment
// padding before every object, vtable written after its table and shared with an
// identical earlier one) in the order the reference's State::save creates objects
// (state.rs:372-392, new_schema_generated.rs:2522-2542), so its output is the
// reference's bytes.
#include <algorithm>
#include <cstring>

#include "psim_host.h"

namespace psim {
namespace host {

namespace {

const char* const kFields[] = {"pos",     "vel",    "acc",       "rot",      "rot_vel", "rot_acc",
                               "forces",  "radii",  "fixity",    "bonds",    "contacts", "material_pointers",
                               "materials", "groups", "scripts", "settings"};
constexpr int kSlots = 16;
constexpr int kSlotScripts = 14;
constexpr int kSlotSettings = 15;
constexpr size_t kSettingsNew = 212;
constexpr size_t kSettingsOld = 204;

bool is_int_field(const std::string& n) {
    return n == "fixity" || n == "bonds" || n == "material_pointers" || n == "groups";
}

template <typename T>
T rd(const uint8_t* buf, size_t len, size_t at) {
    if (at + sizeof(T) > len) throw DemError(DEM_EINVAL, "saved state: read past the end of the buffer");
    T v;
    std::memcpy(&v, buf + at, sizeof(T));
    return v;
}

// Back-to-front builder; bytes live at the end of buf_, `used_` of them.
class Builder {
public:
    size_t used() const { return used_; }

    void align(size_t size, size_t alignment) {
        min_align_ = std::max(min_align_, alignment);
        const size_t pad = (alignment - ((used_ + size) % alignment)) % alignment;
        grow(pad);
        used_ += pad;
        std::memset(head(), 0, pad);
    }
    void prepend(const void* d, size_t n) {
        grow(n);
        used_ += n;
        std::memcpy(head(), d, n);
    }
    size_t push_u32(uint32_t v) {
        align(4, 4);
        prepend(&v, 4);
        return used_;
    }
    size_t push_i32(int32_t v) {
        align(4, 4);
        prepend(&v, 4);
        return used_;
    }
    size_t create_vector(const void* data, size_t count) {
        align(4 * count, 4);
        prepend(data, 4 * count);
        return push_u32((uint32_t)count);
    }
    size_t create_string(const std::string& s) {
        align(s.size() + 1, 4);
        const uint8_t zero = 0;
        prepend(&zero, 1);
        prepend(s.data(), s.size());
        return push_u32((uint32_t)s.size());
    }
    void start_table() {
        fields_.clear();
        table_start_ = used_;
    }
    void add_offset(uint16_t voff, size_t target) {
        align(4, 4);
        const uint32_t rel = (uint32_t)(used_ + 4 - target);
        prepend(&rel, 4);
        fields_.push_back({voff, used_});
    }
    void add_struct(uint16_t voff, const uint8_t* d, size_t n) {
        align(n, 4);
        prepend(d, n);
        fields_.push_back({voff, used_});
    }
    size_t end_table() {
        const size_t obj = push_i32(0);   // soffset to the vtable, patched below
        uint16_t max_voff = 2;
        for (auto& f : fields_) max_voff = std::max(max_voff, f.first);
        const size_t vt_len = (size_t)max_voff + 2;
        std::vector<uint8_t> vt(vt_len, 0);
        const uint16_t hdr[2] = {(uint16_t)vt_len, (uint16_t)(obj - table_start_)};
        std::memcpy(vt.data(), hdr, 4);
        for (auto& f : fields_) {
            const uint16_t o = (uint16_t)(obj - f.second);
            std::memcpy(vt.data() + f.first, &o, 2);
        }
        size_t found = 0;
        for (size_t rev : vtables_) {
            if (std::memcmp(buf_.data() + buf_.size() - rev, vt.data(), vt_len) == 0) {
                found = rev;
                break;
            }
        }
        if (!found) {
            prepend(vt.data(), vt_len);
            found = used_;
            vtables_.push_back(found);
        }
        const int32_t so = (int32_t)found - (int32_t)obj;
        std::memcpy(buf_.data() + buf_.size() - obj, &so, 4);
        return obj;
    }
    std::vector<uint8_t> finish(size_t root) {
        align(4, min_align_);
        align(4, 4);
        const uint32_t rel = (uint32_t)(used_ + 4 - root);
        prepend(&rel, 4);
        return std::vector<uint8_t>(head(), head() + used_);
    }

private:
    uint8_t* head() { return buf_.data() + buf_.size() - used_; }
    void grow(size_t n) {
        if (used_ + n <= buf_.size()) return;
        const size_t cap = std::max(buf_.size() * 2, used_ + n + 1024);
        std::vector<uint8_t> nb(cap);
        std::memcpy(nb.data() + cap - used_, buf_.data() + buf_.size() - used_, used_);
        buf_.swap(nb);
    }
    std::vector<uint8_t> buf_;
    size_t used_ = 0;
    size_t min_align_ = 1;
    std::vector<size_t> vtables_;
    std::vector<std::pair<uint16_t, size_t>> fields_;
    size_t table_start_ = 0;
};

}  // namespace

const std::vector<float>* SavedState::f(const std::string& name) const {
    for (auto& kv : fvec)
        if (kv.first == name) return &kv.second;
    return nullptr;
}

const std::vector<int32_t>* SavedState::i(const std::string& name) const {
    for (auto& kv : ivec)
        if (kv.first == name) return &kv.second;
    return nullptr;
}

SavedState read_state(const uint8_t* buf, size_t len) {
    if (len < 8) throw DemError(DEM_EINVAL, "saved state: buffer too short");
    const size_t root = rd<uint32_t>(buf, len, 0);
    const int64_t vt = (int64_t)root - rd<int32_t>(buf, len, root);
    if (vt < 0 || (size_t)vt + 4 > len) throw DemError(DEM_EINVAL, "saved state: vtable out of range");
    const uint16_t vsz = rd<uint16_t>(buf, len, (size_t)vt);
    const uint16_t tsz = rd<uint16_t>(buf, len, (size_t)vt + 2);
    if (vsz < 4 || vsz % 2 || (size_t)vt + vsz > len) throw DemError(DEM_EINVAL, "saved state: bad vtable size");
    const int nslots = (vsz - 4) / 2;
    std::vector<uint16_t> offs(nslots);
    for (int k = 0; k < nslots; k++) offs[k] = rd<uint16_t>(buf, len, (size_t)vt + 4 + 2 * k);
    // old schema: 15 slots, settings last, the table ending 204 bytes after it
    const bool old = nslots == 15 && tsz != 0 && offs[14] != 0 && (size_t)(tsz - offs[14]) == kSettingsOld;
    SavedState st;
    st.old_schema = old;
    for (int slot = 0; slot < nslots && slot < kSlots; slot++) {
        if (offs[slot] == 0) continue;
        const int field = (old && slot == 14) ? kSlotSettings : slot;
        const std::string name = kFields[field];
        const size_t p = root + offs[slot];
        if (field == kSlotSettings) {
            const size_t size = old ? kSettingsOld : kSettingsNew;
            if (p + size > len) throw DemError(DEM_EINVAL, "saved state: settings out of range");
            st.settings_raw.assign(buf + p, buf + p + size);
            continue;
        }
        const size_t vp = p + rd<uint32_t>(buf, len, p);
        const size_t n = rd<uint32_t>(buf, len, vp);
        if (field == kSlotScripts) {
            if (vp + 4 + n > len) throw DemError(DEM_EINVAL, "saved state: string out of range");
            st.has_scripts = true;
            st.scripts.assign(reinterpret_cast<const char*>(buf + vp + 4), n);
            continue;
        }
        if (vp + 4 + 4 * n > len) throw DemError(DEM_EINVAL, "saved state: vector " + name + " out of range");
        if (is_int_field(name)) {
            std::vector<int32_t> v(n);
            if (n) std::memcpy(v.data(), buf + vp + 4, 4 * n);
            st.ivec.push_back({name, std::move(v)});
        } else {
            std::vector<float> v(n);
            if (n) std::memcpy(v.data(), buf + vp + 4, 4 * n);
            st.fvec.push_back({name, std::move(v)});
        }
    }
    return st;
}

std::vector<uint8_t> write_state(const SavedState& s) {
    if (s.settings_raw.size() != kSettingsNew)
        throw DemError(DEM_EINVAL, "saved state: the settings struct must be 212 bytes");
    Builder b;
    size_t off[kSlots] = {0};
    bool have[kSlots] = {false};
    const char* const order[] = {"pos",    "vel",   "acc",    "rot",      "rot_vel",           "rot_acc",   "forces",
                                 "radii",  "fixity", "bonds", "contacts", "material_pointers", "materials", "groups"};
    for (const char* nm : order) {
        const std::string name = nm;
        int slot = 0;
        while (name != kFields[slot]) slot++;
        if (is_int_field(name)) {
            const std::vector<int32_t>* v = s.i(name);
            if (!v) continue;
            off[slot] = b.create_vector(v->data(), v->size());
        } else {
            const std::vector<float>* v = s.f(name);
            if (!v) continue;
            off[slot] = b.create_vector(v->data(), v->size());
        }
        have[slot] = true;
    }
    if (s.has_scripts) {
        off[kSlotScripts] = b.create_string(s.scripts);
        have[kSlotScripts] = true;
    }
    b.start_table();
    b.add_struct((uint16_t)(4 + 2 * kSlotSettings), s.settings_raw.data(), s.settings_raw.size());
    for (int slot = kSlotSettings - 1; slot >= 0; slot--)
        if (have[slot]) b.add_offset((uint16_t)(4 + 2 * slot), off[slot]);
    const size_t root = b.end_table();
    return b.finish(root);
}

}  // namespace host
}  // namespace psim
