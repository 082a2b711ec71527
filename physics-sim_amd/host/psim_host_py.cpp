// psim_host_py.cpp - Python binding of the C++ host mirror, for the tests: each call
// takes and returns saved-state bytes (the reference's flatbuffer format), so a test
// can compare the C++ mirror with the Python one and with the oracle byte for byte.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <string>
#include <vector>

#include "psim_host.h"

namespace py = pybind11;
using namespace psim::host;

namespace {

std::vector<uint8_t> to_vec(const py::bytes& b) {
    const std::string s = b;
    return std::vector<uint8_t>(s.begin(), s.end());
}

py::bytes to_bytes(const std::vector<uint8_t>& v) { return py::bytes(reinterpret_cast<const char*>(v.data()), v.size()); }

}  // namespace

PYBIND11_MODULE(_psim_host, m) {
    m.doc() = "C++ host mirror (State, Settings, ComputeProg) over libpsim's C ABI";
    py::register_exception<DemError>(m, "DemError");

    // flatbuffer codec: parse and re-serialise
    m.def("codec_roundtrip", [](py::bytes b) {
        const std::vector<uint8_t> v = to_vec(b);
        return to_bytes(write_state(read_state(v.data(), v.size())));
    });
    // State::load then State::save
    m.def("state_roundtrip", [](py::bytes b) {
        const std::vector<uint8_t> v = to_vec(b);
        Settings s;
        State st = State::load(v.data(), v.size(), s);
        return to_bytes(st.save(s));
    });
    // the 29-word uniform and the grids of a saved state's settings
    m.def("collision_settings", [](py::bytes b) {
        const std::vector<uint8_t> v = to_vec(b);
        Settings s;
        State::load(v.data(), v.size(), s);
        const dem_uniform u = s.collision_settings();
        return py::bytes(reinterpret_cast<const char*>(&u), sizeof(u));
    });
    m.def("grids", [](py::bytes b) {
        const std::vector<uint8_t> v = to_vec(b);
        Settings s;
        State st = State::load(v.data(), v.size(), s);
        const dem_grid g = s.grid_info();
        const dem_grid d = ComputeProg::default_grid(s, st);
        return py::make_tuple(py::make_tuple(g.cell_size, g.cell_cap, g.w, g.h),
                              py::make_tuple(d.cell_size, d.cell_cap, d.w, d.h));
    });
    // bond assembly for (i, j) pairs (int32 bytes) over a saved state's positions
    m.def("assemble_bonds", [](py::bytes state, py::bytes pairs, int slots) {
        const std::vector<uint8_t> v = to_vec(state);
        Settings s;
        State st = State::load(v.data(), v.size(), s);
        const std::string p = pairs;
        std::vector<int32_t> pr(p.size() / 4);
        std::memcpy(pr.data(), p.data(), p.size());
        std::vector<int32_t> bonds;
        std::vector<float> contacts;
        assemble_bonds(st, pr, slots, bonds, contacts);
        return py::make_tuple(py::bytes(reinterpret_cast<const char*>(bonds.data()), 4 * bonds.size()),
                              py::bytes(reinterpret_cast<const char*>(contacts.data()), 4 * contacts.size()));
    });
    // load a saved state, run dem_step calls of the given sizes through ComputeProg,
    // download and save: the bytes a C++ client of the library would write
    m.def("run", [](py::bytes b, std::vector<int> calls, bool regen_bonds) {
        const std::vector<uint8_t> v = to_vec(b);
        Settings s;
        State st = State::load(v.data(), v.size(), s);
        ComputeProg::Options o;
        if (regen_bonds) o.bond_capacity = std::max(st.p_count * s.setup.max_bonds, 1);
        ComputeProg prog(s, st, o);
        if (regen_bonds) prog.regen_bonds(s);
        for (int n : calls) prog.step(n);
        prog.update_state();
        double ke = prog.kinetic_energy();
        int64_t broken = prog.broken_bond_count();
        return py::make_tuple(to_bytes(st.save(s)), ke, broken);
    });
}
