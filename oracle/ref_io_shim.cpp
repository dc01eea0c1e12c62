// Thin extern "C" wrapper over the reference's own file READERS, compiled
// straight from the reference tree (see oracle/Makefile) into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: tests/ use it to read the files this project's
// drivers write with the reference's own code.
//
// Wrapped reference code (paths relative to the reference repository root):
//   cpp/include/mipt/density.hpp:253-347   io::DensityMatrixReader (MIPTRHO2 v2)
//   cpp/include/mipt/tmi/resume.hpp:205-333 tmi::resume::load (sim_tmi checkpoint)
#include "mipt/density.hpp"
#include "mipt/tmi/resume.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

static void put_error(const std::exception &e, char *err, int errcap)
{
    if (!err || errcap <= 0) return;
    std::strncpy(err, e.what(), static_cast<std::size_t>(errcap - 1));
    err[errcap - 1] = '\0';
}

extern "C" {

// Header of a MIPTRHO2 file: u32[10] = spatial_dimension, total_qubits,
// kept_qubits, matrix_dimension, periods, realizations, resolution, grid_x,
// grid_y, subsystem_count; *record_count; p[2] = p_min, p_max.  0 on success,
// 1 when the reader throws (message in err).
int ref_density_header(const char *path, std::uint32_t *u32, std::uint64_t *record_count, double *p, char *err,
                       int errcap)
{
    try
    {
        mipt::io::DensityMatrixReader r(path);
        const auto &m = r.metadata();
        const std::uint32_t v[10] = {m.spatial_dimension, m.total_qubits, m.kept_qubits, m.matrix_dimension,
                                     m.periods,           m.realizations, m.resolution,  m.grid_x,
                                     m.grid_y,            m.subsystem_count};
        std::memcpy(u32, v, sizeof(v));
        *record_count = m.record_count;
        p[0] = m.p_min;
        p[1] = m.p_max;
        return 0;
    }
    catch (const std::exception &e)
    {
        put_error(e, err, errcap);
        return 1;
    }
}

// Subsystem list (subsystem_count * kept_qubits) and every record: p_index,
// realization, p and the payload (record_count * subsystem_count * 2 D^2
// doubles), read through DensityMatrixReader::read_record.  Buffers sized from
// ref_density_header.  0 on success, 1 on a reader error.
int ref_density_records(const char *path, std::uint32_t *subsystems, std::uint32_t *p_index,
                        std::uint32_t *realization, double *p, double *rho, char *err, int errcap)
{
    try
    {
        mipt::io::DensityMatrixReader r(path);
        const auto &m = r.metadata();
        std::memcpy(subsystems, m.subsystem_qubits.data(), m.subsystem_qubits.size() * sizeof(std::uint32_t));
        const std::size_t per = mipt::io::payload_value_count(m);
        mipt::io::DensityRecord rec;
        std::size_t i = 0;
        while (r.read_record(rec))
        {
            p_index[i] = rec.p_index;
            realization[i] = rec.realization;
            p[i] = rec.p;
            std::memcpy(rho + i * per, rec.rho_ri.data(), per * sizeof(double));
            ++i;
        }
        return 0;
    }
    catch (const std::exception &e)
    {
        put_error(e, err, errcap);
        return 1;
    }
}

// tmi::resume::load on an existing CSV.  p_texts: the run's p grid as the
// writer renders it, '\n'-separated.  out[5] = rows, dropped_rows,
// kept_bytes, start_p_index, start_realization; flags[2] = name_verified,
// loaded.  0 on success, 1 when the reference refuses (message in err).
int ref_tmi_resume_load(const char *output_path, const char *expected_name, int realizations, int cycle_count,
                        const char *p_texts, std::uint64_t *out, int *flags, char *err, int errcap)
{
    try
    {
        std::vector<std::string> ps;
        std::string cur;
        for (const char *c = p_texts; *c; ++c)
        {
            if (*c == '\n')
            {
                ps.push_back(cur);
                cur.clear();
            }
            else
                cur.push_back(*c);
        }
        if (!cur.empty()) ps.push_back(cur);
        const auto ck = mipt::tmi::resume::load(output_path, expected_name, realizations, cycle_count, ps);
        out[0] = ck.rows;
        out[1] = ck.dropped_rows;
        out[2] = static_cast<std::uint64_t>(ck.kept_bytes);
        out[3] = ck.start_p_index;
        out[4] = static_cast<std::uint64_t>(ck.start_realization);
        flags[0] = ck.name_verified ? 1 : 0;
        flags[1] = ck.loaded ? 1 : 0;
        return 0;
    }
    catch (const std::exception &e)
    {
        put_error(e, err, errcap);
        return 1;
    }
}

}  // extern "C"
