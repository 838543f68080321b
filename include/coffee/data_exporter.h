// data_exporter (src/utils/data_exporter.h:15-100) on one GPU: restart
// snapshots and field output, HDF5 files with the reference's layout, over
// libcoffee_h5.so (include/coffee_h5.h) and the library's sampling kernel
// (coffee_downsample).  Link with -lcoffee_hip -lcoffee_h5.
//
// save_snapshot / load_snapshot (data_exporter.cpp:134-259): the whole guarded
// arrays Ex..Bz, B0x..B0z, P, divE, divB and the scalars "step" (uint32) and
// "time" (Scalar).  write_output / write_field_output (data_exporter.cpp:
// 120-131, 593-643): fld.%05d.h5 in the output directory, each array sampled
// every `downsample` cells and interpolated from its stagger to the cell
// centre, as float; the _cum energy-accounting arrays are reset after output.
#ifndef COFFEE_DATA_EXPORTER_H
#define COFFEE_DATA_EXPORTER_H

#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "coffee/sim_data.h"
#include "coffee/sim_env.h"
#include "coffee_h5.h"
#include "coffee_hip.h"
#include "coffee_io.h"

namespace Coffee {

class data_exporter {
 public:
  data_exporter(sim_environment& env, uint32_t& timestep, const std::string& output_dir = "./Data/")
      : m_env(env), outputDirectory(output_dir) {
    (void)timestep;
    if (outputDirectory.empty() || outputDirectory.back() != '/') outputDirectory += '/';
    std::error_code ec;
    std::filesystem::create_directories(outputDirectory, ec);
    if (std::filesystem::exists("config.toml"))
      std::filesystem::copy_file("config.toml", outputDirectory + "config.toml",
                                 std::filesystem::copy_options::overwrite_existing, ec);
  }
  virtual ~data_exporter() {
    if (m_tmp) (void)hipFree(m_tmp);
  }
  data_exporter(const data_exporter&) = delete;
  data_exporter& operator=(const data_exporter&) = delete;

  const std::string& output_directory() const { return outputDirectory; }

  void write_output(sim_data& data, uint32_t timestep, double time) {
    write_field_output(data, timestep, time);
  }

  void write_field_output(sim_data& data, uint32_t timestep, double time) {
    (void)time;
    char num[32];
    snprintf(num, sizeof(num), "%05u", timestep / (uint32_t)m_env.params().data_interval);
    File f(outputDirectory + "fld." + num + ".h5", true);
    add_grid_output(data.E.data(0), "Ex", data.E.stagger(0), f);
    add_grid_output(data.E.data(1), "Ey", data.E.stagger(1), f);
    add_grid_output(data.E.data(2), "Ez", data.E.stagger(2), f);
    add_grid_output(data.B.data(0), "Bx", data.B.stagger(0), f);
    add_grid_output(data.B.data(1), "By", data.B.stagger(1), f);
    add_grid_output(data.B.data(2), "Bz", data.B.stagger(2), f);
    add_grid_output(data.B0.data(0), "Jx", data.B0.stagger(0), f);
    add_grid_output(data.B0.data(1), "Jy", data.B0.stagger(1), f);
    add_grid_output(data.B0.data(2), "Jz", data.B0.stagger(2), f);
    add_grid_output(data.P, "P", Stagger(0b111), f);
    add_grid_output(data.divB, "divB", Stagger(0b111), f);
    add_grid_output(data.divE, "divE", Stagger(0b111), f);
    add_grid_output(data.dU_EgtB, "dU_EgtB", Stagger(0b111), f);
    add_grid_output(data.dU_Epar, "dU_Epar", Stagger(0b111), f);
    add_grid_output(data.dU_KO, "dU_KO", Stagger(0b111), f);
    add_grid_output(data.dU_EgtB_cum, "dU_EgtB_cum", Stagger(0b111), f);
    add_grid_output(data.dU_Epar_cum, "dU_Epar_cum", Stagger(0b111), f);
    add_grid_output(data.dU_KO_cum, "dU_KO_cum", Stagger(0b111), f);
    data.dU_EgtB_cum.assign_dev(0.0);
    data.dU_Epar_cum.assign_dev(0.0);
    data.dU_KO_cum.assign_dev(0.0);
    data.dU_EgtB_cum.assign(0.0);
    data.dU_Epar_cum.assign(0.0);
    data.dU_KO_cum.assign(0.0);
  }

  void save_snapshot(const std::string& filename, sim_data& data, uint32_t step, Scalar time) {
    data.sync_to_host();
    File f(filename, true);
    for (auto& m : members(data)) f.write(*m.second, m.first, m_env.grid());
    f.scalar_write("step", COFFEE_H5_U32, &step);
    f.scalar_write("time", scalar_type(), &time);
  }

  void load_snapshot(const std::string& filename, sim_data& data, uint32_t& step, Scalar& time) {
    if (!std::filesystem::exists(filename)) {
      std::cout << "Can't find restart file, proceeding without loading it!" << std::endl;
      return;
    }
    File f(filename, false);
    for (auto& m : members(data)) f.read(*m.second, m.first, m_env.grid());
    f.scalar_read("step", COFFEE_H5_U32, &step);
    f.scalar_read("time", scalar_type(), &time);
    data.sync_to_device();
    m_env.send_guard_cells(data);
  }

 protected:
  // One open HDF5 file; closes at scope exit.
  struct File {
    coffee_h5_file* h = nullptr;
    File(const std::string& path, bool create) {
      int rc = create ? coffee_h5_create(path.c_str(), &h) : coffee_h5_open(path.c_str(), 0, &h);
      check(rc);
    }
    ~File() { coffee_h5_close(h); }
    static void check(int rc) {
      if (rc) throw std::runtime_error(std::string("coffee_h5: ") + coffee_h5_last_error());
    }
    static int rank_ext(const Grid& g, unsigned long long ext[3]) {
      int n = g.dim() == 3 ? 3 : 2;
      for (int i = 0; i < 3; i++) ext[i] = (unsigned long long)g.dims[i];
      return n;
    }
    void write(const multi_array<Scalar>& a, const char* name, const Grid& g) {
      unsigned long long ext[3];
      int n = rank_ext(g, ext);
      check(coffee_h5_write_array(h, name, a.host_ptr(), scalar_type(), n, ext));
    }
    void read(multi_array<Scalar>& a, const char* name, const Grid& g) {
      unsigned long long ext[3];
      int n = rank_ext(g, ext);
      check(coffee_h5_read_array(h, name, a.host_ptr(), scalar_type(), n, ext));
    }
    void scalar_write(const char* name, int type, const void* v) {
      check(coffee_h5_write_scalar(h, name, type, v));
    }
    void scalar_read(const char* name, int type, void* v) {
      check(coffee_h5_read_scalar(h, name, type, v));
    }
  };

  static int scalar_type() { return sizeof(Scalar) == 4 ? COFFEE_H5_F32 : COFFEE_H5_F64; }

  static std::vector<std::pair<const char*, multi_array<Scalar>*>> members(sim_data& d) {
    return {{"Ex", &d.E.data(0)}, {"Ey", &d.E.data(1)}, {"Ez", &d.E.data(2)},
            {"Bx", &d.B.data(0)}, {"By", &d.B.data(1)}, {"Bz", &d.B.data(2)},
            {"B0x", &d.B0.data(0)}, {"B0y", &d.B0.data(1)}, {"B0z", &d.B0.data(2)},
            {"P", &d.P}, {"divE", &d.divE}, {"divB", &d.divB}};
  }

  // data_exporter::add_grid_output (data_exporter.cpp:685-705): the sampled
  // array is made on the device and only it is copied to the host.
  void add_grid_output(multi_array<Scalar>& array, const char* name, Stagger st, File& f) {
    const Grid& g = m_env.grid();
    const int d = m_env.params().downsample > 0 ? m_env.params().downsample : 1;
    unsigned long long ext[3];
    for (int i = 0; i < 3; i++) {
      int n = g.reduced_dim(i);
      ext[i] = (unsigned long long)(n > d ? n / d : n);
    }
    const size_t n = (size_t)(ext[0] * ext[1] * ext[2]);
    if (n > m_tmp_n) {
      if (m_tmp) (void)hipFree(m_tmp);
      m_tmp = nullptr;
      hip_check(hipMalloc(&m_tmp, n * sizeof(float)), "data_exporter tmp_grid_data");
      m_tmp_n = n;
    }
    coffee_grid cg = m_env.c_grid();
    const int bits = st[0] | (st[1] << 1) | (st[2] << 2);
    coffee_check(coffee_downsample(&cg, array.dev_ptr(), (int)sizeof(Scalar), bits, d, m_tmp, nullptr),
                 "coffee_downsample");
    m_host.resize(n);
    hip_check(hipMemcpy(m_host.data(), m_tmp, n * sizeof(float), hipMemcpyDeviceToHost),
              "data_exporter sample copy");
    const int rank = ext[2] > 1 ? 3 : 2;
    File::check(coffee_h5_write_array(f.h, name, m_host.data(), COFFEE_H5_F32, rank, ext));
  }

  sim_environment& m_env;
  std::string outputDirectory;
  float* m_tmp = nullptr;
  size_t m_tmp_n = 0;
  std::vector<float> m_host;
};

}  // namespace Coffee

#endif
