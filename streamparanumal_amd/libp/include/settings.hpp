// Named command-line settings (include/settings.hpp:78-139).
// Same semantics as the reference: short and long keys, defaults, allowed
// option lists, "Cannot set setting twice", substring compareSetting.  One
// deliberate change: an unknown flag or a flag without a value raises
// libp::exception instead of looping forever (libs/core/settings.cpp:139-165
// advances only on a match).
#pragma once

#include <map>

#include "comm.hpp"

namespace libp {

class setting_t {
 public:
  std::string shortkey, longkey, name, val, description;
  std::vector<std::string> options;
  int check = 0;

  setting_t() = default;
  setting_t(std::string shortkey_, std::string longkey_, std::string name_, std::string val_,
            std::string description_ = "", std::vector<std::string> options_ = {});

  template <typename T>
  T getVal() const {
    std::stringstream ss(val);
    T arg{};
    ss >> arg;
    return arg;
  }
  void updateVal(const std::string newVal);
  bool compareVal(const std::string token) const;
  std::string toString() const;
  std::string PrintUsage() const;
};

std::ostream &operator<<(std::ostream &os, const setting_t &setting);

class settings_t {
 public:
  comm_t comm;
  std::map<std::string, setting_t> settings;

  settings_t() = default;
  explicit settings_t(comm_t comm_) : comm(comm_) {}

  void newSetting(const std::string shortkey, const std::string longkey, const std::string name,
                  const std::string val, const std::string description = "",
                  const std::vector<std::string> options = {});
  void changeSetting(const std::string name, const std::string newVal);
  // Returns normally after -h/--help too; helpRequested() is then true.
  void parseSettings(const int argc, char **argv);

  template <typename T>
  void getSetting(const std::string name, T &value) const {
    auto search = settings.find(name);
    if (search != settings.end()) value = search->second.getVal<T>();
    else LIBP_FORCE_ABORT("Unable to find setting: [" << name << "]");
  }
  std::string getSetting(const std::string name) const;
  bool compareSetting(const std::string name, const std::string token) const;
  bool hasSetting(const std::string name) const { return settings.count(name) != 0; }
  void report();
  void reportSetting(const std::string name) const;
  void PrintUsage();
  bool helpRequested() const { return help_; }

 private:
  std::vector<std::string> insertOrder_;
  bool help_ = false;
};

}  // namespace libp
