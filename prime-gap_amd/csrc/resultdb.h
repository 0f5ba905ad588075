// resultdb.h -- result persistence for gap_test_hip: the `result` and
// `m_stats` rows the reference's tester records, in an SQLite file.
//
// Reference counterparts: gap_test_stats.save (gap_test_stats.py:217-238:
// REPLACE INTO result; UPDATE m_stats adding the prp counts and test time)
// and load_existing (gap_test_stats.py:145-150: m's of the range already in
// `result` are not tested again).  The tables follow schema.sql:103-154;
// when the file already holds the reference's full schema the statements
// below touch only the columns both share.
//
// The image ships libsqlite3.so.0 without its header, so the few C API
// functions used are declared here (their public, documented signatures).
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <set>
#include <string>
#include <utility>
#include <vector>

extern "C" {
typedef struct sqlite3 sqlite3;
typedef struct sqlite3_stmt sqlite3_stmt;
int sqlite3_open(const char *filename, sqlite3 **db);
int sqlite3_open_v2(const char *filename, sqlite3 **db, int flags, const char *vfs);
int sqlite3_close(sqlite3 *db);
int sqlite3_exec(sqlite3 *db, const char *sql, int (*cb)(void *, int, char **, char **), void *arg, char **err);
int sqlite3_prepare_v2(sqlite3 *db, const char *sql, int nbyte, sqlite3_stmt **stmt, const char **tail);
int sqlite3_bind_int64(sqlite3_stmt *stmt, int idx, long long v);
int sqlite3_bind_double(sqlite3_stmt *stmt, int idx, double v);
int sqlite3_step(sqlite3_stmt *stmt);
int sqlite3_reset(sqlite3_stmt *stmt);
int sqlite3_finalize(sqlite3_stmt *stmt);
long long sqlite3_column_int64(sqlite3_stmt *stmt, int col);
double sqlite3_column_double(sqlite3_stmt *stmt, int col);
int sqlite3_changes(sqlite3 *db);
const char *sqlite3_errmsg(sqlite3 *db);
void sqlite3_free(void *p);
}

namespace pgap {

class ResultDb {
  public:
    ~ResultDb() { close(); }

    bool open(const std::string &path) {
        if (sqlite3_open(path.c_str(), &db_) != 0) return fail("open");
        static const char *ddl =
            "CREATE TABLE IF NOT EXISTS result (P INTEGER, D INTEGER, m INTEGER, next_p INTEGER, prev_p INTEGER, "
            "merit REAL, PRIMARY KEY(P, D, m));"
            "CREATE TABLE IF NOT EXISTS m_stats (P INTEGER, D INTEGER, m INTEGER, next_p INTEGER DEFAULT 0, "
            "prev_p INTEGER DEFAULT 0, merit REAL DEFAULT 0, prp_next INTEGER DEFAULT 0, "
            "prp_prev INTEGER DEFAULT 0, test_time REAL DEFAULT 0, PRIMARY KEY(P, D, m));"
            "BEGIN;";
        if (!exec(ddl)) return false;
        return prepare("REPLACE INTO result(P, D, m, next_p, prev_p, merit) VALUES(?, ?, ?, ?, ?, ?)", &ins_) &&
               prepare("UPDATE m_stats SET next_p = ?, prev_p = ?, merit = ?, prp_next = prp_next + ?, "
                       "prp_prev = prp_prev + ?, test_time = test_time + ? WHERE P = ? AND D = ? AND m = ?",
                       &upd_) &&
               prepare("INSERT INTO m_stats(P, D, m, next_p, prev_p, merit, prp_next, prp_prev, test_time) "
                       "VALUES(?, ?, ?, ?, ?, ?, ?, ?, ?)",
                       &add_);
    }

    // `result` rows of (P, D) with m in [lo, hi]: finished ones (next_p >= 0;
    // 0 marks a side-skipped m) into `done`, partial ones (next_p < 0, tested
    // again whatever their probability -- gap_test.py:331-350) into `partial`.
    bool existing(uint32_t P, uint32_t D, uint64_t lo, uint64_t hi, std::set<uint64_t> &done,
                  std::set<uint64_t> &partial) {
        sqlite3_stmt *q = nullptr;
        if (!prepare("SELECT m, next_p FROM result WHERE P = ? AND D = ? AND m BETWEEN ? AND ?", &q)) return false;
        bind_range(q, P, D, lo, hi);
        int rc;
        while ((rc = sqlite3_step(q)) == 100 /* SQLITE_ROW */)
            (sqlite3_column_int64(q, 1) >= 0 ? done : partial).insert((uint64_t)sqlite3_column_int64(q, 0));
        sqlite3_finalize(q);
        return rc == 101 /* SQLITE_DONE */ || fail("select result");
    }

    // (m, prob_record) of the m_stats rows gap_stats wrote for [lo, hi]
    // (load_probs_only, gap_test_stats.py:170-183), prob_record as float as
    // the reference's array('f') holds it; `merit` reads prob_merit instead.
    // Fails when m_stats has no such column.
    bool probs(uint32_t P, uint32_t D, uint64_t lo, uint64_t hi, std::vector<std::pair<uint64_t, float>> &out,
               bool merit = false) {
        sqlite3_stmt *q = nullptr;
        const std::string sql = std::string("SELECT m, ") + (merit ? "prob_merit" : "prob_record") +
                                " FROM m_stats WHERE P = ? AND D = ? AND m BETWEEN ? AND ? ORDER BY m";
        if (!prepare(sql.c_str(), &q)) return false;
        bind_range(q, P, D, lo, hi);
        int rc;
        while ((rc = sqlite3_step(q)) == 100 /* SQLITE_ROW */)
            out.emplace_back((uint64_t)sqlite3_column_int64(q, 0), (float)sqlite3_column_double(q, 1));
        sqlite3_finalize(q);
        return rc == 101 /* SQLITE_DONE */ || fail("select m_stats");
    }

    // merit is stored rounded to 4 places, as gap_test_stats.save does.
    bool save(uint32_t P, uint32_t D, uint64_t m, int next_p, int prev_p, double merit, long long n_tests,
              long long p_tests, double test_time) {
        const double mr = (double)(long long)(merit * 10000.0 + (merit >= 0 ? 0.5 : -0.5)) / 10000.0;
        sqlite3_reset(ins_);
        sqlite3_bind_int64(ins_, 1, P);
        sqlite3_bind_int64(ins_, 2, D);
        sqlite3_bind_int64(ins_, 3, (long long)m);
        sqlite3_bind_int64(ins_, 4, next_p);
        sqlite3_bind_int64(ins_, 5, prev_p);
        sqlite3_bind_double(ins_, 6, mr);
        if (sqlite3_step(ins_) != 101) return fail("result");
        sqlite3_reset(upd_);
        sqlite3_bind_int64(upd_, 1, next_p);
        sqlite3_bind_int64(upd_, 2, prev_p);
        sqlite3_bind_double(upd_, 3, mr);
        sqlite3_bind_int64(upd_, 4, n_tests);
        sqlite3_bind_int64(upd_, 5, p_tests);
        sqlite3_bind_double(upd_, 6, test_time);
        sqlite3_bind_int64(upd_, 7, P);
        sqlite3_bind_int64(upd_, 8, D);
        sqlite3_bind_int64(upd_, 9, (long long)m);
        if (sqlite3_step(upd_) != 101) return fail("m_stats update");
        if (sqlite3_changes(db_) == 0) {  // no gap_stats row for this m: add one
            sqlite3_reset(add_);
            sqlite3_bind_int64(add_, 1, P);
            sqlite3_bind_int64(add_, 2, D);
            sqlite3_bind_int64(add_, 3, (long long)m);
            sqlite3_bind_int64(add_, 4, next_p);
            sqlite3_bind_int64(add_, 5, prev_p);
            sqlite3_bind_double(add_, 6, mr);
            sqlite3_bind_int64(add_, 7, n_tests);
            sqlite3_bind_int64(add_, 8, p_tests);
            sqlite3_bind_double(add_, 9, test_time);
            if (sqlite3_step(add_) != 101) return fail("m_stats insert");
        }
        if (++pending_ >= 1000) {
            pending_ = 0;
            return exec("COMMIT; BEGIN;");
        }
        return true;
    }

    void close() {
        for (sqlite3_stmt *s : {ins_, upd_, add_})
            if (s) sqlite3_finalize(s);
        ins_ = upd_ = add_ = nullptr;
        if (db_) {
            exec("COMMIT;");
            sqlite3_close(db_);
            db_ = nullptr;
        }
    }

    const std::string &error() const { return err_; }

  private:
    sqlite3 *db_ = nullptr;
    sqlite3_stmt *ins_ = nullptr, *upd_ = nullptr, *add_ = nullptr;
    int pending_ = 0;
    std::string err_;

    bool fail(const char *what) {
        err_ = std::string(what) + ": " + (db_ ? sqlite3_errmsg(db_) : "no database");
        return false;
    }
    bool exec(const char *sql) {
        char *msg = nullptr;
        if (sqlite3_exec(db_, sql, nullptr, nullptr, &msg) != 0) {
            err_ = msg ? msg : "sqlite error";
            sqlite3_free(msg);
            return false;
        }
        return true;
    }
    static void bind_range(sqlite3_stmt *q, uint32_t P, uint32_t D, uint64_t lo, uint64_t hi) {
        sqlite3_bind_int64(q, 1, P);
        sqlite3_bind_int64(q, 2, D);
        sqlite3_bind_int64(q, 3, (long long)lo);
        sqlite3_bind_int64(q, 4, (long long)hi);
    }
    bool prepare(const char *sql, sqlite3_stmt **s) {
        return sqlite3_prepare_v2(db_, sql, -1, s, nullptr) == 0 || fail("prepare");
    }
};

// The record gap sizes of load_records (gap_test_stats.py:116-142), read
// from a prime-gap list file (table gaps(gapsize, merit, ...)), opened read
// only: every gap whose merit at ln N = log_n would beat its record
// (gapsize > merit * log_n - 0.001) up to merit 35, and, when 30 ln N
// exceeds 113326 (the smallest gap size missing from the list), every even
// size from 113326 to 30 ln N absent from the list.  Sorted ascending.
inline bool load_record_gaps(const std::string &path, double log_n, std::vector<int64_t> &out, std::string &err) {
    sqlite3 *db = nullptr;
    auto fail = [&](const char *what) {
        err = std::string(what) + ": " + (db ? sqlite3_errmsg(db) : "cannot open");
        if (db) sqlite3_close(db);
        return false;
    };
    if (sqlite3_open_v2(path.c_str(), &db, 0x1 /* SQLITE_OPEN_READONLY */, nullptr) != 0) return fail("open");
    sqlite3_stmt *q = nullptr;
    if (sqlite3_prepare_v2(db, "SELECT gapsize FROM gaps WHERE gapsize > merit * ? - 0.001 AND gapsize < 35 * ?",
                           -1, &q, nullptr) != 0)
        return fail("select gaps");
    sqlite3_bind_double(q, 1, log_n);
    sqlite3_bind_double(q, 2, log_n);
    int rc;
    while ((rc = sqlite3_step(q)) == 100) out.push_back(sqlite3_column_int64(q, 0));
    sqlite3_finalize(q);
    if (rc != 101) return fail("select gaps");
    const int64_t smallest_missing = 113326, largest_likely = (int64_t)(log_n * 30);
    if (smallest_missing < largest_likely) {
        std::set<int64_t> listed;
        if (sqlite3_prepare_v2(db, "SELECT gapsize FROM gaps WHERE gapsize BETWEEN ? AND ?", -1, &q, nullptr) != 0)
            return fail("select gaps");
        sqlite3_bind_int64(q, 1, smallest_missing);
        sqlite3_bind_int64(q, 2, largest_likely);
        while ((rc = sqlite3_step(q)) == 100) listed.insert(sqlite3_column_int64(q, 0));
        sqlite3_finalize(q);
        if (rc != 101) return fail("select gaps");
        for (int64_t g = smallest_missing; g <= largest_likely; g += 2)
            if (!listed.count(g)) out.push_back(g);
    }
    sqlite3_close(db);
    std::sort(out.begin(), out.end());
    return true;
}

}  // namespace pgap
