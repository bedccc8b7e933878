// validation.h — the host program's validation mode (SIM_VALIDATION,
// PandemicThrust/PandemicThrust/PandemicSim.h:50-58; the checks of
// debug.cu).  It writes debug.txt (progress lines, the per-day block and an
// "ERROR: ..." line for every failed check, in the reference's three assert
// formats, PandemicSim.cu:1436-1475), debug_infected.csv (header only, as
// the reference writes it), debug_contacts.csv (validateContacts_wholeDay,
// debug.cu:138-442) and debug_actions.csv (debug_validateActions,
// debug.cu:847-950) from the device trace of pt_sim_trace.
//
// Observation point.  pt_sim_step runs a whole day (recovery sweep,
// infected list, contacts, infections), so the per-day checks look at the
// state after the day's infections are written.  The infection-status check
// of debug.cu:710-786 is restated for that point: an agent is active on day
// d when a strain is infected with day <= d (infections made on day d carry
// day d + 1, PandemicSim.cu:3048-3067), and an infected strain's profile day
// d - day lies in [-1, CULMINATION_PERIOD).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pandemic/pt_abi.h"

class Validator {
public:
    ~Validator() { close(); }
    bool open(const std::string& out_dir);
    void print(const char* msg);                                   // debug_print
    void check(bool ok, const char* msg);                          // debug_assert(bool, msg)
    void check(bool ok, const char* msg, int idx);                 // debug_assert(bool, msg, idx)
    void expect(const char* msg, int expected, int actual);        // debug_assert(msg, expected, actual)

    // debug_validatePeopleSetup (debug.cu:664-708) after pt_sim_setup.
    void setup(pt_sim* sim);
    // After pt_sim_step of day `day` with infected > 0: the infection-status
    // check, the day block, the errand schedule, contacts and actions.
    void day(pt_sim* sim, int day, int infected);
    void close();

private:
    void infection_status(pt_sim* sim, int day, int infected);
    void errand_schedule(bool weekend);
    void contacts(int day, bool weekend);
    void actions(int day);

    FILE* f_debug_ = nullptr;
    FILE* f_infected_ = nullptr;
    FILE* f_contacts_ = nullptr;
    FILE* f_actions_ = nullptr;
    int N_ = 0, H_ = 0, W_ = 0;
    std::vector<uint8_t> age_;
    std::vector<int32_t> household_, workplace_, type_off_;
    // the day's state and trace
    std::vector<uint8_t> st_p_, st_s_, dy_p_, dy_s_, gn_p_, gn_s_;
    std::vector<int32_t> infected_, errands_;
    std::vector<int32_t> t_infector_, t_victim_, t_loc_;
    std::vector<uint8_t> t_type_, t_action_;
    int slots_ = 0;  // contacts per infector today
};
