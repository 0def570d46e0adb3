// dcns/model.h -- value types of the reference's hot path, with the same
// public interface (C++14), implemented for this library.
//
//   Seed                  model/Seed.h:13-43
//   PairedSimilarFragment model/PairedSimilarFragment.h:14-68
//   Scorei                model/Score.h (Scorei), model/Score.cpp:105-143
//   Matrix                model/Matrix.h:10-21
#ifndef DCNS_MODEL_H
#define DCNS_MODEL_H

#include <cstdint>
#include <string>
#include <vector>

// Four 0-based coordinates of an ungapped-or-gapped seed.  operator< is the
// reference's comparator: two seeds sharing their start pair, or their end
// pair, are equivalent; otherwise order by (start1, start2).  It is not a
// strict weak order, and std::set<Seed> behaviour under it is part of the
// results (which seeds survive, in which order).
class Seed {
  private:
    int32_t start1;
    int32_t end1;
    int32_t start2;
    int32_t end2;

  public:
    Seed(const int32_t &_start1, const int32_t &_end1, const int32_t &_start2, const int32_t &_end2)
        : start1(_start1), end1(_end1), start2(_start2), end2(_end2) {}
    int32_t getStart1() const { return start1; }
    void setStart1(const int32_t &v) { start1 = v; }
    int32_t getEnd1() const { return end1; }
    void setEnd1(const int32_t &v) { end1 = v; }
    int32_t getStart2() const { return start2; }
    void setStart2(const int32_t &v) { start2 = v; }
    int32_t getEnd2() const { return end2; }
    void setEnd2(const int32_t &v) { end2 = v; }
    bool operator<(const Seed &c) const {
        if (start1 == c.start1 && start2 == c.start2) return false;
        if (end1 == c.end1 && end2 == c.end2) return false;
        if (start1 < c.start1) return true;
        if (start1 == c.start1) return start2 < c.start2;
        return false;
    }
};

// A reported conserved fragment: 1-based inclusive coordinates, score,
// cigar ops (len << 4 | op; 0 M, 1 I, 2 D), p-value and e-value.
class PairedSimilarFragment {
  private:
    uint32_t start1 = 0, end1 = 0, start2 = 0, end2 = 0, score = 0;
    std::vector<uint32_t> cigar;
    double pValue = 0, eValue = 0;
    std::string alignment1, alignment2;

  public:
    PairedSimilarFragment() = default;
    PairedSimilarFragment(uint32_t _start1, uint32_t _end1, uint32_t _start2, uint32_t _end2, uint32_t _score,
                          const std::vector<uint32_t> &_cigar, const double &_pValue, const double &_eValue)
        : start1(_start1), end1(_end1), start2(_start2), end2(_end2), score(_score), cigar(_cigar), pValue(_pValue),
          eValue(_eValue) {}
    uint32_t getStart1() const { return start1; }
    void setStart1(uint32_t v) { start1 = v; }
    uint32_t getEnd1() const { return end1; }
    void setEnd1(uint32_t v) { end1 = v; }
    uint32_t getStart2() const { return start2; }
    void setStart2(uint32_t v) { start2 = v; }
    uint32_t getEnd2() const { return end2; }
    void setEnd2(uint32_t v) { end2 = v; }
    uint32_t getScore() const { return score; }
    void setScore(uint32_t v) { score = v; }
    uint32_t getLength() const { return end2 - start2 + 1; }
    const std::vector<uint32_t> &getCigar() const { return cigar; }
    void setCigar(const std::vector<uint32_t> &v) { cigar = v; }
    double getPValue() const { return pValue; }
    void setPValue(double v) { pValue = v; }
    double getEValue() const { return eValue; }
    void setEValue(double v) { eValue = v; }
    const std::string &getAlignment1() const { return alignment1; }
    void setAlignment1(const std::string &v) { alignment1 = v; }
    const std::string &getAlignment2() const { return alignment2; }
    void setAlignment2(const std::string &v) { alignment2 = v; }
};

// 5x5 substitution table over codes A0 C1 N2 G3 T4; N scores 0.
class Scorei {
  private:
    int8_t m[5][5];

  public:
    Scorei(const int8_t &matchingScore, const int8_t &mismatchingPenalty) {
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j)
                m[i][j] = (i == 2 || j == 2) ? 0 : (i == j ? matchingScore : mismatchingPenalty);
    }
    const int8_t &getScore(const int8_t &a, const int8_t &b) const { return m[a][b]; }
};

// Traceback scratch of the reference's alignment routines.  The device
// kernels keep their own band-compressed traceback, so this type only
// carries the dimensions for signature compatibility.
class Matrix {
  private:
    int32_t length1, length2;

  public:
    Matrix(int32_t _length1, int32_t _length2) : length1(_length1), length2(_length2) {}
    int32_t rows() const { return length1; }
    int32_t cols() const { return length2; }
};

#endif  // DCNS_MODEL_H
