// icrar::MeasurementSet over in-memory columns.
//
// The accessor contract the calibration path uses
// (src/icrar/leap-accelerate/ms/MeasurementSet.h:54-224, MeasurementSet.cc:33-294):
// station, baseline, row, timestep, channel and polarisation counts, epochs,
// flagged / short / filtered baselines, coordinates and visibilities.  casacore
// is not part of this library: the columns come from the caller (a table
// reader, a synthetic generator, or device memory already holding DATA).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <memory>
#include <optional>
#include <set>
#include <sstream>
#include <vector>

#include "icrar/leap-accelerate/exception/exception.h"
#include "leap_hip.h"

namespace icrar {

// Columns of a measurement set, in MS row order (time-major, nbaselines rows per epoch).
struct MeasurementSetColumns {
    uint32_t total_antennas = 0;        // rows of the ANTENNA table
    uint32_t num_rows = 0;              // main-table rows
    uint32_t num_channels = 0;          // SPECTRAL_WINDOW NUM_CHAN[0]
    uint32_t num_pols = 0;              // POLARIZATION NUM_CORR[0]
    double ref_frequency_hz = 0.0;      // SPECTRAL_WINDOW REF_FREQUENCY[0]
    double chan_width_hz = 0.0;         // SPECTRAL_WINDOW CHAN_WIDTH[0][0]
    double phase_centre_ra_rad = 0.0;   // FIELD PHASE_DIR[0]
    double phase_centre_dec_rad = 0.0;
    std::vector<int32_t> antenna1, antenna2;  // num_rows each
    std::vector<double> time;                 // num_rows
    std::vector<uint8_t> flag_c0p0;           // FLAG[row][0][0], num_rows (or at least one epoch)
    std::vector<double> uvw;                  // num_rows x 3
    // DATA as [row][channel][pol] complex: host memory, or device memory of the calibrator's device.
    const void* data = nullptr;
    leap_vis_dtype data_dtype = LEAP_VIS_C64;
    bool data_on_device = false;
    const double* uvw_device = nullptr;       // optional device copy of uvw
    std::shared_ptr<const void> data_owner;   // keeps `data` alive (e.g. a file mapping), may be empty
};

class MeasurementSet {
    MeasurementSetColumns m_cols;
    std::set<int32_t> m_antennas;
    uint32_t m_stations = 0;
    bool m_readAutocorrelations = true;

public:
    // MeasurementSet.cc:33-62: stations are the unique antennas unless overridden.
    MeasurementSet(MeasurementSetColumns cols, std::optional<int> overrideNStations, bool readAutocorrelations)
        : m_cols(std::move(cols)), m_readAutocorrelations(readAutocorrelations) {
        m_antennas = CalculateUniqueAntennas();
        if (overrideNStations) {
            m_stations = uint32_t(*overrideNStations);
        } else {
            m_stations = uint32_t(m_antennas.size());
        }
        Validate();
    }

    const MeasurementSetColumns& GetColumns() const { return m_cols; }

    uint32_t GetTotalAntennas() const { return m_cols.total_antennas; }
    uint32_t GetNumStations() const { return m_stations; }
    uint32_t GetNumBaselines() const { return GetNumBaselines(m_readAutocorrelations); }
    uint32_t GetNumBaselines(bool useAutocorrelations) const {
        const uint32_t n = GetNumStations();
        return useAutocorrelations ? n * (n + 1) / 2 : n * (n - 1) / 2;  // MeasurementSet.cc:154-167
    }
    uint32_t GetNumPols() const { return m_cols.num_pols; }
    uint32_t GetNumChannels() const { return m_cols.num_channels; }
    uint32_t GetNumRows() const { return m_cols.num_rows; }
    uint32_t GetNumTimesteps() const { return GetNumRows() / GetNumBaselines(); }

    // MeasurementSet.cc:120-130: TIME of the first row of every epoch.
    std::vector<double> GetEpochs() const {
        std::vector<double> r;
        for (uint32_t i = 0; i < GetNumTimesteps(); ++i) r.push_back(m_cols.time[size_t(i) * GetNumBaselines()]);
        return r;
    }

    // MeasurementSet.cc:181-197: FLAG of the first epoch, channel 0, pol 0.
    std::vector<bool> GetFlaggedBaselines() const {
        std::vector<bool> r(GetNumBaselines());
        for (uint32_t i = 0; i < GetNumBaselines(); ++i) r[i] = m_cols.flag_c0p0[i] != 0;
        return r;
    }
    uint32_t GetNumFlaggedBaselines() const { return count(GetFlaggedBaselines()); }

    // MeasurementSet.cc:205-228 as the reference's release build evaluates it:
    // the slicer keeps u only and the (1 x rows) result is indexed (i, 0), (i, 1),
    // i.e. u of rows i and i + 1.
    std::vector<bool> GetShortBaselines(double minimumBaselineThreshold = 0.0) const {
        std::vector<bool> r(GetNumBaselines(), false);
        if (minimumBaselineThreshold > 0.0) {
            for (uint32_t i = 0; i < GetNumBaselines(); ++i) {
                const double x = m_cols.uvw[size_t(i) * 3];
                const double y = (i + 1 < GetNumRows()) ? m_cols.uvw[size_t(i + 1) * 3] : 0.0;
                if (std::sqrt(x * x + y * y) < minimumBaselineThreshold) r[i] = true;
            }
        }
        return r;
    }
    uint32_t GetNumShortBaselines(double t = 0.0) const { return count(GetShortBaselines(t)); }
    std::vector<bool> GetFilteredBaselines(double t = 0.0) const {
        std::vector<bool> a = GetFlaggedBaselines(), b = GetShortBaselines(t);
        for (size_t i = 0; i < a.size(); ++i) a[i] = a[i] || b[i];
        return a;
    }
    uint32_t GetNumFilteredBaselines(double t = 0.0) const { return count(GetFilteredBaselines(t)); }

    // The reference's unique-antenna count reads ANTENNA1 twice (MeasurementSet.cc:287-294).
    std::set<int32_t> CalculateUniqueAntennas() const {
        return std::set<int32_t>(m_cols.antenna1.begin(), m_cols.antenna1.end());
    }

    // The C view handed to the calibrator library.
    leap_ms_view View() const {
        leap_ms_view v{};
        v.num_stations = GetNumStations();
        v.total_antennas = GetTotalAntennas();
        v.num_baselines = GetNumBaselines();
        v.num_rows = GetNumRows();
        v.num_channels = GetNumChannels();
        v.num_pols = GetNumPols();
        v.freq_start_hz = m_cols.ref_frequency_hz;
        v.freq_inc_hz = m_cols.chan_width_hz;
        v.phase_centre_ra_rad = m_cols.phase_centre_ra_rad;
        v.phase_centre_dec_rad = m_cols.phase_centre_dec_rad;
        v.antenna1 = m_cols.antenna1.data();
        v.antenna2 = m_cols.antenna2.data();
        v.time = m_cols.time.data();
        v.flag = m_cols.flag_c0p0.data();
        v.uvw = m_cols.uvw.data();
        v.vis = m_cols.data;
        v.vis_dtype = m_cols.data_dtype;
        v.vis_on_device = m_cols.data_on_device ? 1 : 0;
        v.uvw_device = m_cols.uvw_device;
        return v;
    }

private:
    static uint32_t count(const std::vector<bool>& v) { return uint32_t(std::count(v.begin(), v.end(), true)); }

    // MeasurementSet.cc:64-103
    void Validate() const {
        const size_t nrows = m_cols.num_rows;
        if (m_cols.antenna1.size() < nrows || m_cols.antenna2.size() < nrows || m_cols.time.size() < nrows ||
            m_cols.uvw.size() < 3 * nrows)
            throw icrar::exception("measurement set columns shorter than the row count", __FILE__, __LINE__);
        if (nrows == 0) throw icrar::exception("measurement set has no rows", __FILE__, __LINE__);
        const double epoch = m_cols.time[0];
        const auto epochRows = std::count(m_cols.time.begin(), m_cols.time.begin() + nrows, epoch);
        if (uint32_t(epochRows) != GetNumBaselines())
            throw icrar::exception("visibilities at first epoch does not match number of baselines", __FILE__, __LINE__);
        if (GetNumRows() < GetNumBaselines()) {
            std::stringstream ss;
            ss << "invalid number of rows, expected >=" << GetNumBaselines() << ", got " << GetNumRows();
            throw icrar::exception(ss.str(), __FILE__, __LINE__);
        }
        if (GetNumBaselines() == 0 || GetNumRows() % GetNumBaselines() != 0)
            throw icrar::exception("number of rows not an integer multiple of baselines", __FILE__, __LINE__);
        if (m_cols.flag_c0p0.size() < GetNumBaselines())
            throw icrar::exception("flag column shorter than one epoch", __FILE__, __LINE__);
    }
};

}  // namespace icrar
