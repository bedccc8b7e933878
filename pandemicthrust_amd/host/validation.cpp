// validation.cpp — see validation.h.  Citations are relative to
// PandemicThrust/PandemicThrust/ of the reference.
#include "validation.h"

#include <algorithm>
#include <cstdlib>

namespace {

// Codes of simParameters.h:23-36, :59-64, :84-90 and PandemicSim.h:297-302.
constexpr int kAgeAdult = 5;
constexpr int kSusceptible = 1, kInfected = 2, kRecovered = 0;
constexpr int kProfiles = 6, kCulmination = 10;
constexpr int kDayNone = 255, kGenNone = 255;
constexpr int kPreschool = 3, kUniversity = 7, kAfterschool = 8, kFirstErrandRow = 9;
constexpr int kWeekendErrands = 3, kWeekdayErrands = 2, kWeekendHours = 10;
enum { CT_NONE = 0, CT_WORKPLACE = 1, CT_SCHOOL = 2, CT_ERRAND = 3, CT_AFTERSCHOOL = 4, CT_HOME = 5 };

// CHILD_AGE_SCHOOLTYPE_LOOKUP_HOST (PandemicSim.cu:336-340)
const int kChildSchoolType[5] = {3, 4, 5, 6, 7};

// lookup_contact_type (PandemicSim.cu:1557-1576)
const char* contact_type_name(int t) {
    switch (t) {
        case CT_NONE: return "CONTACT_TYPE_NONE";
        case CT_WORKPLACE: return "CONTACT_TYPE_WORKPLACE";
        case CT_SCHOOL: return "CONTACT_TYPE_SCHOOL";
        case CT_ERRAND: return "CONTACT_TYPE_ERRAND";
        case CT_AFTERSCHOOL: return "CONTACT_TYPE_AFTERSCHOOL";
        case CT_HOME: return "CONTACT_TYPE_HOME";
        default: return "BAD_CONTACT_TYPE_NUM";
    }
}

// action_type_to_string (PandemicSim.cu:1497-1512): 1 seasonal, 2 pandemic.
const char* action_name(int a) {
    switch (a) {
        case 0: return "NONE";
        case 2: return "PAND";
        case 1: return "SEAS";
        case 3: return "BOTH";
        default: return "????";
    }
}

bool is_infected(int status) { return status >= kInfected; }

template <typename T>
void copy_array(pt_sim* sim, int which, std::vector<T>* out) {
    const int64_t bytes = pt_sim_array_bytes(sim, which);
    if (bytes < 0) {
        fprintf(stderr, "pandemicthrust: validation: array %d: %s\n", which, pt_last_error());
        exit(1);
    }
    out->resize((size_t)bytes / sizeof(T));
    if (bytes > 0 && pt_sim_copy_array(sim, which, out->data(), bytes) != PT_OK) {
        fprintf(stderr, "pandemicthrust: validation: array %d: %s\n", which, pt_last_error());
        exit(1);
    }
}

}  // namespace

// The output streams of logging_openOutputStreams (PandemicSim.cu:210-262).
bool Validator::open(const std::string& out_dir) {
    auto path = [&](const char* n) { return out_dir + "/" + n; };
    f_infected_ = fopen(path("debug_infected.csv").c_str(), "w");
    f_contacts_ = fopen(path("debug_contacts.csv").c_str(), "w");
    f_actions_ = fopen(path("debug_actions.csv").c_str(), "w");
    f_debug_ = fopen(path("debug.txt").c_str(), "w");
    if (!f_infected_ || !f_contacts_ || !f_actions_ || !f_debug_) return false;
    fprintf(f_infected_, "current_day, i, idx, status_p, day_p, gen_p, status_s, day_s, gen_s\n");
    fprintf(f_contacts_, "current_day, i, infector_idx, victim_idx, contact_type, contact_loc, infector_loc, "
                         "victim_loc, locs_matched\n");
    fprintf(f_actions_, "current_day, i, infector, victim, action_type, action_type_string\n");
    return true;
}

void Validator::close() {
    for (FILE** f : {&f_infected_, &f_contacts_, &f_actions_, &f_debug_})
        if (*f) {
            fclose(*f);
            *f = nullptr;
        }
}

void Validator::print(const char* msg) {
    fprintf(f_debug_, "%s\n", msg);
    fflush(f_debug_);
}

void Validator::check(bool ok, const char* msg) {
    if (!ok) fprintf(f_debug_, "ERROR: %s\n", msg);
}

void Validator::check(bool ok, const char* msg, int idx) {
    if (!ok) fprintf(f_debug_, "ERROR: %s index: %d\n", msg, idx);
}

void Validator::expect(const char* msg, int expected, int actual) {
    if (expected != actual) fprintf(f_debug_, "ERROR: %s expected: %d actual: %d\n", msg, expected, actual);
}

// debug_validatePeopleSetup (debug.cu:664-708).  The adult and child totals
// of the household scan are not exported, so the last two checks of the
// reference are replaced by the household offsets covering every agent.
void Validator::setup(pt_sim* sim) {
    int64_t sc[8];
    if (pt_sim_get_scalars(sim, sc) != PT_OK) {
        fprintf(stderr, "pandemicthrust: validation: %s\n", pt_last_error());
        exit(1);
    }
    N_ = (int)sc[0];
    H_ = (int)sc[1];
    W_ = (int)sc[2];
    copy_array(sim, PT_AGE, &age_);
    copy_array(sim, PT_HOUSEHOLD, &household_);
    copy_array(sim, PT_WORKPLACE, &workplace_);
    copy_array(sim, PT_LOCATION_TYPE_OFFSETS, &type_off_);
    std::vector<int32_t> hh_off;
    copy_array(sim, PT_HH_OFFSETS, &hh_off);
    check(std::is_sorted(household_.begin(), household_.end()), "household indexes are not monotonically increasing");
    for (int i = 0; i < N_; ++i) {
        const int a = age_[(size_t)i], hh = household_[(size_t)i], wp = workplace_[(size_t)i];
        check(hh >= 0 && hh < H_, "invalid household ID assigned to person", i);
        check(wp >= 0 && wp < W_, "invalid workplace ID assigned to person", i);
        if (a == kAgeAdult) continue;
        if (a >= 0 && a < kAgeAdult) {
            const int t = kChildSchoolType[a];
            check(wp >= type_off_[(size_t)t] && wp < type_off_[(size_t)t + 1], "school is not valid for age, person", i);
        } else {
            check(false, "invalid age code for person", i);
        }
    }
    expect("household offsets do not cover the population", N_, hh_off.empty() ? -1 : hh_off.back());
}

// debug_validateInfectionStatus (debug.cu:710-786) at the observation point
// described in validation.h.
void Validator::infection_status(pt_sim* sim, int day, int infected) {
    copy_array(sim, PT_STATUS_P, &st_p_);
    copy_array(sim, PT_STATUS_S, &st_s_);
    copy_array(sim, PT_DAY_P, &dy_p_);
    copy_array(sim, PT_DAY_S, &dy_s_);
    copy_array(sim, PT_GEN_P, &gn_p_);
    copy_array(sim, PT_GEN_S, &gn_s_);
    copy_array(sim, PT_INFECTED, &infected_);
    expect("infected list length", infected, (int)infected_.size());
    int count = 0;
    for (int i = 0; i < N_; ++i) {
        const int sp = st_p_[(size_t)i], ss = st_s_[(size_t)i];
        const int dp = dy_p_[(size_t)i], ds = dy_s_[(size_t)i];
        const int gp = gn_p_[(size_t)i], gs = gn_s_[(size_t)i];
        const bool active_p = is_infected(sp) && sp - kInfected < kProfiles && dp <= day;
        const bool active_s = is_infected(ss) && ss - kInfected < kProfiles && ds <= day;
        const bool listed = std::binary_search(infected_.begin(), infected_.end(), i);
        if (active_p || active_s) {
            ++count;
            check(listed, "active infection, but index not in infected_indexes, person", i);
        } else {
            check(!listed, "no active infection, but index in infected_indexes, person", i);
        }
        if (sp == kSusceptible) {
            check(dp == kDayNone, "susceptible to pandemic but day is set, person", i);
            check(gp == kGenNone, "susceptible to pandemic but generation is set, person", i);
        } else if (sp == kRecovered) {
            check(dp < day, "recovered from pandemic but day is invalid");
            check(gp < day, "recovered from pandemic but gen is invalid");
        } else {
            check(is_infected(sp) && sp - kInfected < kProfiles, "invalid pandemic status/profile code, person index",
                  i);
            const int profile_day = day - dp;
            check(profile_day >= -1 && profile_day < kCulmination, "pandemic day is outside of valid range");
        }
        if (ss == kSusceptible) {
            check(ds == kDayNone, "susceptible to seasonal but day is set, person", i);
            check(gs == kGenNone, "susceptible to seasonal but generation is set, person", i);
        } else if (ss == kRecovered) {
            check(ds < day, "recovered from seasonal but day is invalid");
            check(gs < day, "recovered from seasonal but gen is invalid");
        } else {
            check(is_infected(ss) && ss - kInfected < kProfiles, "invalid seasonal status/profile code, person");
            const int profile_day = day - ds;
            check(profile_day >= -1 && profile_day < kCulmination, "seasonal day is outside of valid range");
        }
    }
    check(count == infected, "validation count of infected disagrees with sim");
}

// debug_validateErrandSchedule (debug.cu:522-566).
void Validator::errand_schedule(bool weekend) {
    const int per = weekend ? kWeekendErrands : kWeekdayErrands;
    const int first_row = kFirstErrandRow;
    for (int i = 0; i < N_; ++i) {
        const bool adult = age_[(size_t)i] == kAgeAdult;
        for (int e = 0; e < per; ++e) {
            const int32_t errand = weekend ? errands_[(size_t)i * kWeekendErrands + e] : errands_[(size_t)e * N_ + i];
            const int loc = errand % W_, hour = errand / W_;
            if (weekend)
                check(hour >= 0 && hour < kWeekendHours, "scheduled weekend errand has invalid hour, person", i);
            else
                expect("scheduled weekday errand has unexpected hour", e, hour);
            if (adult || weekend)
                check(loc >= type_off_[(size_t)first_row] && loc < W_, "person location is not valid for errand PDF, person",
                      i);
            else
                check(loc >= type_off_[kAfterschool] && loc < type_off_[kAfterschool + 1],
                      "person afterschool destination is not afterschool location type, person", i);
        }
    }
}

// validateContacts_wholeDay (debug.cu:138-442): one row per contact slot.
void Validator::contacts(int day, bool weekend) {
    const int n = (int)t_victim_.size();
    for (int k = 0; k < n; ++k) {
        const int pos = k / slots_;
        const int infected_index = infected_[(size_t)pos];
        const int type = t_type_[(size_t)k];
        const int infector = t_infector_[(size_t)k];
        int victim = t_victim_[(size_t)k];
        const int location = t_loc_[(size_t)k];
        const int infector_age = age_[(size_t)infected_index];
        int locs_matched = 0;
        bool abort_check = false;
        if (infector != infected_index) {
            print("error: contact_infector improperly set");
            abort_check = true;
        }
        if (infector < 0 || infector >= N_) abort_check = true;
        if (victim < 0 || victim >= N_) abort_check = true;
        if (type == CT_NONE) {
            if (location == -1) {
                expect("nulled contact, but victim is not null", -1, victim);
            } else {
                expect("infector wasn't alone at location, but failed to make contacts, loc_count", 1, victim);
                victim = -1;
            }
            locs_matched = 1;
            abort_check = true;
        }
        int infector_loc = -1, victim_loc = -1;
        if (!abort_check) {
            const int victim_age = age_[(size_t)victim];
            if (type == CT_WORKPLACE) {
                infector_loc = workplace_[(size_t)infector];
                victim_loc = workplace_[(size_t)victim];
                locs_matched = infector_loc == victim_loc;
                check(infector_age == kAgeAdult, "work-type contact but infector is a child, infector", infected_index);
                check(location == infector_loc,
                      "workplace contact: contact_location does not match infector location, person", infected_index);
                check(infector_loc >= 0 && infector_loc < W_,
                      "infector location out of bounds in workplace contact, infector", infected_index);
                check(victim_loc >= 0 && victim_loc < W_, "victim location out of bounds in workplace contact, infector",
                      infected_index);
                check(locs_matched, "workplace contact between people with different workplaces, infector",
                      infected_index);
                check(!weekend, "work contact during weekend, infector", infected_index);
            } else if (type == CT_SCHOOL) {
                infector_loc = workplace_[(size_t)infector];
                victim_loc = workplace_[(size_t)victim];
                locs_matched = infector_loc == victim_loc;
                const int lo = type_off_[kPreschool], hi = type_off_[kUniversity + 1];
                check(infector_age != kAgeAdult, "school-type contact but infector is an adult, infector",
                      infected_index);
                check(location == infector_loc,
                      "school contact: contact does not take place at infector's school, person", infected_index);
                check(infector_loc >= lo && infector_loc < hi,
                      "infector loc for school contact is not school locationtype, infector", infected_index);
                check(victim_loc >= lo && victim_loc < hi,
                      "victim loc for school contact is not school locationtype, infector", infected_index);
                check(locs_matched, "school contact between people with different schools, infector", infected_index);
                check(!weekend, "school contact during weekend, infector", infected_index);
            } else if (type == CT_ERRAND && weekend) {
                for (int e = 0; e < kWeekendErrands; ++e) {
                    const int32_t a = errands_[(size_t)infected_index * kWeekendErrands + e];
                    if (a == location) infector_loc = a % W_;
                    const int32_t b = errands_[(size_t)victim * kWeekendErrands + e];
                    if (b == location) victim_loc = b % W_;
                }
                locs_matched = infector_loc >= 0 && victim_loc >= 0 && infector_loc == victim_loc;
                const int lo = type_off_[kFirstErrandRow];
                if (infector_loc != -1)
                    check(infector_loc >= lo && infector_loc < W_,
                          "infector location is not valid for weekend errand PDF, infector", infected_index);
                if (victim_loc != -1)
                    check(victim_loc >= lo && victim_loc < W_,
                          "victim location is not valid for weekend errand PDF, infector", infected_index);
                check(locs_matched, "infector and victim errand destinations do not match, infector", infected_index);
            } else if (type == CT_ERRAND) {
                check(infector_age == kAgeAdult, "weekday errand contact type but infector is a child, infector",
                      infected_index);
                check(victim_age == kAgeAdult, "weekday errand contact type but victim is a child, victim", victim);
                for (int hour = 0; hour < kWeekdayErrands; ++hour) {
                    const int a = errands_[(size_t)hour * N_ + infected_index] % W_;
                    const int b = errands_[(size_t)hour * N_ + victim] % W_;
                    if (a == b) {
                        locs_matched = 1;
                        infector_loc = a;
                        victim_loc = b;
                        break;
                    }
                }
                const int lo = type_off_[kFirstErrandRow];
                if (infector_loc != -1)
                    check(infector_loc >= lo && infector_loc < W_,
                          "infector location is not valid for weekday errand PDF, infector", infected_index);
                if (victim_loc != -1)
                    check(victim_loc >= lo && victim_loc < W_,
                          "victim location is not valid for weekday errand PDF, infector", infected_index);
                check(locs_matched, "infector and victim errand destinations do not match, infector", infected_index);
            } else if (type == CT_AFTERSCHOOL) {
                infector_loc = errands_[(size_t)infected_index] % W_;
                victim_loc = errands_[(size_t)victim] % W_;
                locs_matched = infector_loc == victim_loc;
                const int lo = type_off_[kAfterschool], hi = type_off_[kAfterschool + 1];
                check(infector_age != kAgeAdult, "afterschool contact but infector is an adult, infector",
                      infected_index);
                check(victim_age != kAgeAdult, "afterschool contact but victim is an adult, victim", victim);
                check(location == infector_loc, "contact is not at infector's scheduled afterschool, infector",
                      infected_index);
                check(infector_loc >= lo && infector_loc < hi,
                      "infector afterschool destination is not afterschool location type, infector", infected_index);
                check(victim_loc >= lo && victim_loc < hi,
                      "victim afterschool destination is not afterschool location type, infector", infected_index);
                check(locs_matched, "infector and victim afterschool destinations do not match, infector",
                      infected_index);
                check(!weekend, "afterschool contact on weekend, infector", infected_index);
            } else if (type == CT_HOME) {
                infector_loc = household_[(size_t)infected_index];
                victim_loc = household_[(size_t)victim];
                locs_matched = infector_loc == victim_loc;
                check(location == infector_loc, "household contact not at infector's household, infector",
                      infected_index);
                check(infector_loc >= 0 && infector_loc < H_, "infector household index out of range, infector",
                      infected_index);
                check(victim_loc >= 0 && victim_loc < H_, "victim household index out of range, infector",
                      infected_index);
                check(locs_matched, "household contact between people with different households, infector",
                      infected_index);
            } else {
                print("error: invalid contact type");
            }
        }
        fprintf(f_contacts_, "%d,%d,%d,%d,%s,%d,%d,%d,%d\n", day, k, infector, victim, contact_type_name(type), location,
                infector_loc, victim_loc, locs_matched);
    }
    fflush(f_contacts_);
}

// debug_validateActions (debug.cu:847-950).  The trace carries no uniforms
// or probabilities, so the reference's probability checks become the
// property they guard: a contact infects with a strain only if its infector
// carries that strain, infectious since an earlier day.
void Validator::actions(int day) {
    const int n = (int)t_victim_.size();
    for (int k = 0; k < n; ++k) {
        const int infector = t_infector_[(size_t)k], victim = t_victim_[(size_t)k];
        const int action = t_action_[(size_t)k];
        if (action > 3) print("invalid action code");
        if (action & 2)
            check(is_infected(st_p_[(size_t)infector]) && dy_p_[(size_t)infector] <= day,
                  "infector does not have pandemic, but chance is > 0, infector ", infector);
        if (action & 1)
            check(is_infected(st_s_[(size_t)infector]) && dy_s_[(size_t)infector] <= day,
                  "infector does not have seasonal, but chance is > 0, infector", infector);
        fprintf(f_actions_, "%d,%d,%d,%d,%d,%s\n", day, k, infector, victim, action, action_name(action));
    }
    fflush(f_actions_);
}

void Validator::day(pt_sim* sim, int day, int infected) {
    infection_status(sim, day, infected);
    fprintf(f_debug_, "\n\n---------------------\nday %d\ninfected: %d\n---------------------\n\n", day, infected);
    fflush(f_debug_);
    copy_array(sim, PT_ERRANDS, &errands_);
    const bool weekend = errands_.size() == (size_t)N_ * kWeekendErrands;  // is_weekend (PandemicSim.h:284)
    slots_ = weekend ? 5 : 8;
    errand_schedule(weekend);
    const int64_t n = pt_sim_trace_size(sim);
    if (n < 0 || n != (int64_t)infected * slots_) {
        fprintf(stderr, "pandemicthrust: validation: trace size %lld for %d infected\n", (long long)n, infected);
        exit(1);
    }
    t_infector_.resize((size_t)n);
    t_victim_.resize((size_t)n);
    t_loc_.resize((size_t)n);
    t_type_.resize((size_t)n);
    t_action_.resize((size_t)n);
    if (n > 0 && pt_sim_trace(sim, t_infector_.data(), t_victim_.data(), t_type_.data(), t_loc_.data(),
                              t_action_.data(), n) < 0) {
        fprintf(stderr, "pandemicthrust: validation: %s\n", pt_last_error());
        exit(1);
    }
    contacts(day, weekend);
    actions(day);
    fflush(f_debug_);
}
