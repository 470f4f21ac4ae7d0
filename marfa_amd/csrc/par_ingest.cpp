// par_ingest.cpp -- HITRAN .par text -> the 36-byte direct-access line records the LBL
// path reads (src/LBL.f90:43).  Host code only (no device work): a C++ restatement of the
// reference's preprocessing program app/processParFile.f90:32-185 behind the C ABI.
//
// Each 160-column .par line is read with the reference's edit descriptors
// (processParFile.f90:104):
//   (I2, A1, F12.6, 2E10.3, 2F5.4, F10.4, F4.2, F8.6, 4A15, 6I1, 6I2, A1, F7.1, F7.2)
// under Fortran's rules for formatted numeric input: blanks in a field are ignored (an
// all-blank field reads 0), a field without a decimal point takes d implied decimals, and
// an exponent may be written as E/D followed by a signed integer or as a bare sign.  S and
// E'' are read in FP64 and rounded to FP32 (processParFile.f90:166-167), the other real
// fields directly to FP32 (one correct rounding of the decimal string each).
//
// Record r (1-based) receives input line r: nu (f8), S, gamma_air, gamma_self, E'',
// n_T (f4), jointMolIso (i4) = N_MOL_ISO(iso, mol)*100 + mol, delta (f4)
// (processParFile.f90:169-177).  A line with an unknown isotopologue character is skipped
// with its record left unwritten (a hole of zeros, :159-164).  The reference's read loop
// (:149-181) makes one more pass when the end of the file is met; what that pass writes
// depends on the Fortran runtime (the reference's own fixture test/testDB.02 has no record
// after the last line; built with amdflang the program repeats the last line as record
// N + 1).  This converter writes none, as the fixture shows.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/marfa.h"

namespace {

// N_MOL_ISO(20, 12) (processParFile.f90:80-103): internal isotopologue number of HITRAN
// isotopologue `iso` (1..20) of molecule `mol` (1..12); 0 where the molecule has fewer
constexpr int N_ISO = 20, N_MOL = 12;
constexpr int ISO_COUNT[N_MOL] = {9, 13, 18, 5, 9, 4, 6, 3, 2, 1, 2, 2};
constexpr int ISO_FIRST[N_MOL] = {1, 10, 23, 41, 46, 55, 59, 65, 68, 70, 71, 73};

int mol_iso_number(int iso, int mol) {
    if (iso < 1 || iso > N_ISO || mol < 1 || mol > N_MOL) return -1;
    return iso <= ISO_COUNT[mol - 1] ? ISO_FIRST[mol - 1] + iso - 1 : 0;
}

// convert_ISO (processParFile.f90:227-270): '1'..'9' -> 1..9, '0' -> 10, 'A'..'I' -> 11..19
int convert_iso(char c) {
    if (c >= '1' && c <= '9') return c - '0';
    if (c == '0') return 10;
    if (c >= 'A' && c <= 'I') return 11 + (c - 'A');
    return 99;
}

// Fortran formatted input of one numeric field (blanks ignored).  kind: 'I' integer,
// 'R' real with d implied decimals.  Returns false on characters Fortran rejects.  The
// value is returned as a normalized decimal string mant "e" exp (correctly rounded later).
bool fortran_field(const char *f, int w, char kind, int d, std::string &dec, long &ival) {
    std::string s;
    for (int i = 0; i < w; i++)
        if (f[i] != ' ' && f[i] != '\0' && f[i] != '\n' && f[i] != '\r') s += f[i];
    if (s.empty()) {
        dec = "0";
        ival = 0;
        return true;
    }
    size_t p = 0;
    bool neg = false;
    if (s[p] == '+' || s[p] == '-') neg = s[p++] == '-';
    if (kind == 'I') {
        if (p >= s.size()) return false;
        long v = 0;
        for (; p < s.size(); p++) {
            if (s[p] < '0' || s[p] > '9') return false;
            v = v * 10 + (s[p] - '0');
        }
        ival = neg ? -v : v;
        return true;
    }
    std::string digits;
    int point = -1, ndig = 0;
    for (; p < s.size(); p++) {
        const char c = s[p];
        if (c >= '0' && c <= '9') {
            digits += c;
            ndig++;
        } else if (c == '.' && point < 0) {
            point = ndig;
        } else {
            break;
        }
    }
    if (ndig == 0) return false;
    long ex = 0;
    if (p < s.size()) {  // exponent: E/D/Q [sign] digits, or sign digits
        char c = s[p];
        if (c == 'E' || c == 'e' || c == 'D' || c == 'd' || c == 'Q' || c == 'q') c = s[++p];
        if (p >= s.size()) return false;
        bool eneg = false;
        if (s[p] == '+' || s[p] == '-') eneg = s[p++] == '-';
        if (p >= s.size()) return false;
        for (; p < s.size(); p++) {
            if (s[p] < '0' || s[p] > '9') return false;
            ex = ex * 10 + (s[p] - '0');
            if (ex > 100000) return false;
        }
        if (eneg) ex = -ex;
    }
    // value = 0.digits scaled: digits * 10^(ex - fractional digits)
    const long frac = point < 0 ? d : ndig - point;
    dec = (neg ? "-" : "") + digits + "e" + std::to_string(ex - frac);
    return true;
}

struct ParLine {
    int mol;
    char iso;
    double nu;
    double S, E;  // read in FP64 (rounded to FP32 on output)
    float gair, gself, nT, delta;
};

// one .par line; false on a field Fortran's read would reject (iostat > 0)
bool read_par_line(const char *line, ParLine &o) {
    char buf[161];
    size_t n = strlen(line);
    while (n > 0 && (line[n - 1] == '\n' || line[n - 1] == '\r')) n--;
    // short lines are padded with blanks (PAD='YES')
    memset(buf, ' ', 160);
    memcpy(buf, line, n < 160 ? n : 160);
    buf[160] = '\0';
    std::string d;
    long iv;
    // I2, A1
    if (!fortran_field(buf + 0, 2, 'I', 0, d, iv)) return false;
    o.mol = (int)iv;
    o.iso = buf[2];
    struct RF {
        int off, w, d;
    };
    // F12.6 nu, E10.3 S, E10.3 A, F5.4 gair, F5.4 gself, F10.4 E'', F4.2 n, F8.6 delta
    const RF rf[8] = {{3, 12, 6}, {15, 10, 3}, {25, 10, 3}, {35, 5, 4}, {40, 5, 4}, {45, 10, 4}, {55, 4, 2}, {59, 8, 6}};
    std::string v[8];
    for (int k = 0; k < 8; k++)
        if (!fortran_field(buf + rf[k].off, rf[k].w, 'R', rf[k].d, v[k], iv)) return false;
    // 4A15 (67..126): text; 6I1 (127..132), 6I2 (133..144): integers; A1; F7.1, F7.2
    for (int k = 0; k < 6; k++)
        if (!fortran_field(buf + 127 + k, 1, 'I', 0, d, iv)) return false;
    for (int k = 0; k < 6; k++)
        if (!fortran_field(buf + 133 + 2 * k, 2, 'I', 0, d, iv)) return false;
    if (!fortran_field(buf + 146, 7, 'R', 1, d, iv)) return false;
    if (!fortran_field(buf + 153, 7, 'R', 2, d, iv)) return false;
    o.nu = strtod(v[0].c_str(), nullptr);
    o.S = strtod(v[1].c_str(), nullptr);
    o.gair = strtof(v[3].c_str(), nullptr);
    o.gself = strtof(v[4].c_str(), nullptr);
    o.E = strtod(v[5].c_str(), nullptr);
    o.nT = strtof(v[6].c_str(), nullptr);
    o.delta = strtof(v[7].c_str(), nullptr);
    return true;
}

void pack_record(const ParLine &l, int32_t joint, unsigned char out[36]) {
    const float S = (float)l.S, E = (float)l.E;
    memcpy(out + 0, &l.nu, 8);
    memcpy(out + 8, &S, 4);
    memcpy(out + 12, &l.gair, 4);
    memcpy(out + 16, &l.gself, 4);
    memcpy(out + 20, &E, 4);
    memcpy(out + 24, &l.nT, 4);
    memcpy(out + 28, &joint, 4);
    memcpy(out + 32, &l.delta, 4);
}

}  // namespace

extern "C" int marfa_species_code(const char *species, char code[3]) {
    static const char *names[N_MOL] = {"H2O", "CO2", "O3", "N2O", "CO", "CH4", "O2", "NO", "SO2", "NO2", "NH3", "HNO3"};
    if (!code) return MARFA_E_ARG;
    strcpy(code, "00");
    if (!species) return MARFA_E_SPECIES;
    // trim(adjustl(species)) (processParFile.f90:196)
    const char *b = species;
    while (*b == ' ') b++;
    size_t n = strlen(b);
    while (n > 0 && b[n - 1] == ' ') n--;
    for (int m = 0; m < N_MOL; m++)
        if (strlen(names[m]) == n && strncmp(names[m], b, n) == 0) {
            snprintf(code, 3, "%02d", m + 1);
            return MARFA_OK;
        }
    return MARFA_E_SPECIES;
}

extern "C" int marfa_par_convert(const char *species, const char *par_path, const char *db_path, int64_t *nrec) {
    if (nrec) *nrec = 0;
    char code[3];
    if (!par_path || !db_path) return MARFA_E_ARG;
    const int rc = marfa_species_code(species, code);
    if (rc) return rc;
    FILE *in = fopen(par_path, "r");
    if (!in) return MARFA_E_ARG;
    FILE *out = fopen(db_path, "wb");
    if (!out) {
        fclose(in);
        return MARFA_E_ARG;
    }
    int status = MARFA_OK;
    int64_t counter = 0, written_max = 0;
    unsigned char rec[36];
    char *line = nullptr;
    size_t cap = 0;
    while (getline(&line, &cap, in) >= 0) {
        counter++;
        ParLine l;
        if (!read_par_line(line, l)) {
            status = MARFA_E_PARREAD;
            break;
        }
        const int is = convert_iso(l.iso);
        if (is == 99) {  // processParFile.f90:159-164: warning, record left unwritten
            fprintf(stderr, "WARNING: unknown ISO value: %c skipping ...\n", l.iso);
            continue;
        }
        const int num = mol_iso_number(is, l.mol);
        if (num < 0) {  // outside N_MOL_ISO: the reference indexes out of bounds
            status = MARFA_E_PARREAD;
            break;
        }
        pack_record(l, (int32_t)(num * 100 + l.mol), rec);
        if (fseek(out, (long)((counter - 1) * 36), SEEK_SET) != 0 || fwrite(rec, 1, 36, out) != 36) {
            status = MARFA_E_ARG;
            break;
        }
        written_max = counter;
    }
    free(line);
    fclose(in);
    if (fclose(out) != 0 && status == MARFA_OK) status = MARFA_E_ARG;
    if (nrec) *nrec = written_max;
    return status;
}
